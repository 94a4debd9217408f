"""In-tree build of the HIP engine library for gfx950.

``build()`` always recompiles ``fluidsim_amd/libfluidsim_amd.so`` with hipcc
(cross-compiling works without a GPU).  The engine is split over several
translation units that compile in parallel.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libfluidsim_amd.so")

SOURCES = [
    "fsx_engine.hip",
    "fsx_api_ns2d.hip",
    "fsx_api_ns3d.hip",
    "fsx_api_strat.hip",
    "fsx_api_diag.hip",
    "fsx_api_cnn.hip",
]

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the engine
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(8, n, int(os.environ.get("MAX_JOBS", "8"))))


def build(verbose=False, extra_flags=()):
    objs, procs, pending = [], [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        pending.append((src, cmd))
        objs.append(obj)
    failed = []
    jobs = _jobs()
    while pending or procs:
        while pending and len(procs) < jobs:
            src, cmd = pending.pop(0)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                 stderr=subprocess.STDOUT, text=True)))
        src, proc = procs.pop(0)
        out, _ = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    for obj in objs:
        os.remove(obj)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
