"""Summarise rocprofv3 result databases (rocpd SQLite) as text.

    python -m fluidsim_amd.rocpd_summary kernels DB [DB_PMC ...]
    python -m fluidsim_amd.rocpd_summary traffic FETCH_DB WRITE_DB COPY_MIB [KEY_PREFIX]

``kernels``: per kernel name, calls, average / total duration (from the
kernel trace); with PMC databases, the average counter value per dispatch of
each kernel is added (FETCH_SIZE / WRITE_SIZE are in KiB per dispatch).
"""

import re
import sqlite3
import sys
from collections import defaultdict


def short_name(name):
    """Readable kernel name: template arguments of the pass kernels kept."""
    m = re.search(r"fsx::(k_\w+)<([^>]*(?:<[^>]*>)?[^>]*)>", name)
    if m:
        args = m.group(2)
        args = args.replace("fsx::", "").replace("float", "f32").replace("double", "f64")
        return f"{m.group(1)}<{args}>"
    m = re.search(r"fsx::(k_\w+)", name)
    if m:
        return m.group(1)
    return name[:80]


def kernel_table(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels").fetchall()
    agg = defaultdict(list)
    for name, dur in rows:
        agg[short_name(name)].append(dur)
    return agg


def pmc_table(db):
    con = sqlite3.connect(db)
    try:
        rows = con.execute(
            "select k.name, p.counter_name, p.value from pmc_events p "
            "join kernels k on p.event_id = k.event_id"
        ).fetchall()
    except sqlite3.Error:
        cols = [r[1] for r in con.execute("pragma table_info('counters_collection')")]
        name_col = "kernel_name" if "kernel_name" in cols else "name"
        rows = con.execute(
            f"select {name_col}, counter_name, value from counters_collection"
        ).fetchall()
    agg = defaultdict(lambda: defaultdict(list))
    for name, counter, value in rows:
        agg[short_name(name)][counter].append(value)
    return agg


def dispatch_values(db, counter):
    """(kernel name, duration ns, value, grid size) per dispatch."""
    con = sqlite3.connect(db)
    return con.execute(
        "select kernel_name, duration, value, grid_size from counters_collection "
        "where counter_name = ? order by dispatch_id", (counter,)
    ).fetchall()


def traffic(fetch_db, write_db, copy_mib, prefix=""):
    """HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE (KiB per dispatch),
    each scaled by what the counter reads for the device-to-device copy of
    COPY_MIB MiB: factor = true KiB / counted KiB. The calibration copy is
    the first copyBuffer dispatch that counts at least 1 MiB (bench.py
    times its 1 GiB device copy before anything else runs; later copies,
    e.g. of whole state fields, have other sizes).  A kernel's value is the
    median over its dispatches (a dispatch whose counters come back short
    does not drag it down); the range over the dispatches is printed."""
    import json

    true_kib = copy_mib * 1024.0
    out = {}
    factors = {}
    for counter, db in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        rows = dispatch_values(db, counter)
        copies = [r for r in rows if "copyBuffer" in r[0]]
        if not copies:
            raise SystemExit(f"no copyBuffer dispatch in {db}")
        large = [r for r in copies if r[2] >= 1024]
        if not large:
            raise SystemExit(f"no calibration copy in {db}")
        cal = large[0]
        factors[counter] = true_kib / cal[2]
        agg = defaultdict(list)
        for name, dur, val, grid in rows:
            agg[short_name(name)].append(val)
        for name, vals in agg.items():
            v = sorted(vals)
            med = v[len(v) // 2] if len(v) % 2 else 0.5 * (v[len(v) // 2 - 1] + v[len(v) // 2])
            out.setdefault(name, {})[counter] = med
            out[name][counter + "_range"] = (v[0], v[-1], len(v))
    print(f"calibration: {copy_mib} MiB device copy; FETCH_SIZE factor "
          f"{factors['FETCH_SIZE']:.3f}, WRITE_SIZE factor {factors['WRITE_SIZE']:.3f}")
    print(f"{'kernel':72s} {'FETCH KiB':>14s} {'WRITE KiB':>14s} {'HBM MB/launch':>14s}"
          "  dispatches, FETCH and WRITE min-max KiB")
    res = {}
    key = lambda kv: -(kv[1].get("FETCH_SIZE", 0) + kv[1].get("WRITE_SIZE", 0))
    for name, c in sorted(out.items(), key=key):
        if "FETCH_SIZE" not in c or "WRITE_SIZE" not in c:
            continue
        b = 1024.0 * (c["FETCH_SIZE"] * factors["FETCH_SIZE"]
                      + c["WRITE_SIZE"] * factors["WRITE_SIZE"])
        res[name] = b
        fr, wr = c["FETCH_SIZE_range"], c["WRITE_SIZE_range"]
        print(f"{name[:72]:72s} {c['FETCH_SIZE']:14.1f} {c['WRITE_SIZE']:14.1f} {b / 1e6:14.1f}"
              f"  {fr[2]}, {fr[0]:.0f}-{fr[1]:.0f}, {wr[0]:.0f}-{wr[1]:.0f}")
    if prefix:
        print(json.dumps({prefix + k: v for k, v in res.items()}))
    return res


def main(argv):
    if len(argv) >= 4 and argv[0] == "traffic":
        traffic(argv[1], argv[2], float(argv[3]), argv[4] if len(argv) > 4 else "")
        return 0
    if len(argv) < 2 or argv[0] != "kernels":
        print(__doc__)
        return 2
    kt = kernel_table(argv[1])
    pmcs = [pmc_table(db) for db in argv[2:]]
    total = sum(sum(v) for v in kt.values())
    counters = sorted({c for p in pmcs for k in p.values() for c in k})
    hdr = f"{'kernel':60s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}"
    for c in counters:
        hdr += f" {c + '/disp':>18s}"
    print(hdr)
    for name, durs in sorted(kt.items(), key=lambda kv: -sum(kv[1])):
        line = (f"{name[:60]:60s} {len(durs):6d} {sum(durs) / len(durs) / 1e3:10.1f} "
                f"{sum(durs) / 1e6:10.2f} {100 * sum(durs) / total:6.1f}")
        for c in counters:
            vals = [v for p in pmcs for v in p.get(name, {}).get(c, [])]
            line += f" {(sum(vals) / len(vals)) if vals else float('nan'):18.1f}"
        print(line)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
