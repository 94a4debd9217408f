"""GPU tests of the batched ns2d ensembles (fluidsim_amd.ensemble).

Each run of an ensemble is an ns2d simulation at its own fixed dt and
viscosities; the oracle advances every run alone (FP64 NS2D of
oracle/solvers.py, the restatement of fluidsim/solvers/ns2d/solver.py and
the IF-RK4 of pseudo_spect.py:798-984) from the same initial state.  Bars:
FP64 per-run relative L2 <= 1e-12; the per-run dt equals fluidsim's initial
CFL step (base.py:320-366) to 1e-13; graph replay and eager launches give
identical states; FP32 runs within 1e-4 of the FP64 oracle.
"""

import json

import numpy as np
import pytest

from fluidsim_amd import _lib
from oracle.operators import Oper2D
from oracle.solvers import NS2D
from fluidsim_amd.synthetic import ns2d_noise_rot_fft

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def make_runs(spec, seed0=0):
    """spec: list of (n, count); nu_2 log-uniform in [1e-4, 1e-2]."""
    rng = np.random.default_rng(seed0)
    runs = []
    k = 0
    for n, count in spec:
        for _ in range(count):
            runs.append({"n": n, "nu_2": float(10 ** rng.uniform(-4, -2)),
                         "rot_fft": ns2d_noise_rot_fft(n, n, seed=seed0 + k, kmax=n // 6)})
            k += 1
    return runs


def cfl_dt(oper, rot, cfl=0.5, dt_max=0.2):
    ux, uy = oper.vecfft_from_rotfft(rot)
    tmp = np.abs(oper.ifft(ux)).max() / oper.deltax + np.abs(oper.ifft(uy)).max() / oper.deltay
    return min(cfl / tmp, dt_max)


def test_ensemble_fp64_vs_oracle(tmp_path):
    from fluidsim_amd.ensemble import EnsembleNS2D

    runs = make_runs([(16, 3), (32, 2), (64, 2)], seed0=10)
    runs[1]["nu_4"] = 1e-6
    print("case ensemble FP64 16^2 x3, 32^2 x2, 64^2 x2, 10 steps", flush=True)
    ens = EnsembleNS2D(runs, precision="float64", tune_cache=str(tmp_path / "tune.json"))
    S0 = [ens.get_state(i) for i in range(len(runs))]
    ens.step(10)
    ens.check_nan()
    for i, r in enumerate(runs):
        n = r["n"]
        oper = Oper2D(n, n)
        assert ens.runs[i]["deltat"] == pytest.approx(cfl_dt(oper, S0[i]), rel=1e-13)
        orc = NS2D(oper, nu_2=r["nu_2"], nu_4=r.get("nu_4", 0.0),
                   deltat0=ens.runs[i]["deltat"])
        orc.state_spect[0] = S0[i]
        orc.run(10)
        err = rel_l2(ens.get_state(i), orc.state_spect[0])
        print(f"run {i} n={n}: rel L2 = {err:.3e}")
        assert err <= 1e-12
    ens.close()


def test_graph_replay_equals_eager(tmp_path):
    from fluidsim_amd.ensemble import EnsembleNS2D

    runs = make_runs([(32, 3), (64, 2)], seed0=20)
    cache = str(tmp_path / "tune.json")
    a = EnsembleNS2D(runs, precision="float32", tune_cache=cache)
    b = EnsembleNS2D(runs, precision="float32", tune_cache=cache)
    a.step(6, graph=False)
    b.step(6, graph=True)
    for i in range(len(runs)):
        assert np.array_equal(a.get_state(i), b.get_state(i)), i
    a.close()
    b.close()


def test_autotune_cache_and_tile_independence(tmp_path):
    from fluidsim_amd.ensemble import EnsembleNS2D

    runs = make_runs([(64, 4)], seed0=30)
    cache = tmp_path / "tune.json"
    a = EnsembleNS2D(runs, precision="float32", tune_cache=str(cache))
    table = json.loads(cache.read_text())
    (key, entry), = table.items()
    assert "|ns2d-v2|64|4|float32|4" in key and "candidates (lines/row)" in entry
    assert len(entry["candidates (lines/row)"]) == 3 * 8  # lines 0, 1..64 x 3 row passes
    assert a.tuning[64]["lines"] == entry["lines"] and a.tuning[64]["row"] == entry["row"]
    # the cached choice is reused
    b = EnsembleNS2D(runs, precision="float32", tune_cache=str(cache))
    assert (b.groups[0].lines, b.groups[0].row) == (a.groups[0].lines, a.groups[0].row)
    # the tile width does not change results; the row pass changes FP32
    # rounding only
    c = EnsembleNS2D(runs, precision="float32", autotune=False)
    _lib.call("fsx_batch_set_row", c.groups[0].handle, a.groups[0].row)
    for e in (a, c):
        e.step(4, graph=False)
    for i in range(len(runs)):
        assert np.array_equal(a.get_state(i), c.get_state(i)), i
    for e in (a, b, c):
        e.close()


@pytest.mark.parametrize("row", [0, 1, 2])
def test_row_passes_vs_oracle(row, tmp_path):
    """Each row pass the autotuner may pick (register rows with radix-8 or
    radix-4 stages, the LDS row pass): FP32 runs of two sizes against the
    FP64 oracle after 20 RK4 steps (1e-4, as the cfg-3 subset), and FP64
    runs (row passes 0 and 2; 1 is FP32-only) within 1e-12."""
    from fluidsim_amd.ensemble import EnsembleNS2D

    print(f"case ensemble row pass {row}", flush=True)
    for prec, tol in (("float32", 1e-4), ("float64", 1e-12)):
        if prec == "float64" and row == 1:
            continue
        runs = make_runs([(64, 2), (128, 1)], seed0=50)
        ens = EnsembleNS2D(runs, precision=prec, autotune=False)
        for g in ens.groups:
            _lib.call("fsx_batch_set_row", g.handle, row)
        S0 = [ens.get_state(i).astype(np.complex128) for i in range(len(runs))]
        ens.step(20)
        ens.check_nan()
        for i, r in enumerate(runs):
            n = r["n"]
            orc = NS2D(Oper2D(n, n), nu_2=r["nu_2"], deltat0=ens.runs[i]["deltat"])
            orc.state_spect[0] = S0[i]
            orc.run(20)
            err = rel_l2(ens.get_state(i), orc.state_spect[0])
            print(f"row pass {row} {prec} run {i} n={n}: rel L2 = {err:.3e}")
            assert err <= tol
        ens.close()


def test_cfg3_subset_fp32_error(tmp_path):
    """A slice of the 512-run ensemble (128^2, 256^2, 512^2, nu_2
    log-uniform, fixed CFL dt), 50 RK4 steps, FP32 vs the FP64 oracle."""
    from fluidsim_amd.ensemble import EnsembleNS2D

    runs = make_runs([(128, 3), (256, 2), (512, 1)], seed0=40)
    print("case ensemble FP32 128^2 x3, 256^2 x2, 512^2 x1, 50 steps", flush=True)
    ens = EnsembleNS2D(runs, precision="float32", tune_cache=str(tmp_path / "t.json"))
    S0 = [ens.get_state(i).astype(np.complex128) for i in range(len(runs))]
    ens.step(50)
    ens.check_nan()
    for i in (0, 3, 5):
        r = runs[i]
        n = r["n"]
        orc = NS2D(Oper2D(n, n), nu_2=r["nu_2"], deltat0=ens.runs[i]["deltat"])
        orc.state_spect[0] = S0[i]
        orc.run(50)
        err = rel_l2(ens.get_state(i), orc.state_spect[0])
        print(f"run {i} n={n} nu_2={r['nu_2']:.2e}: FP32 rel L2 = {err:.3e}")
        assert err < 1e-4
    ens.close()


def test_rk2_ensemble(tmp_path):
    from fluidsim_amd.ensemble import EnsembleNS2D

    runs = make_runs([(32, 2)], seed0=50)
    ens = EnsembleNS2D(runs, precision="float64", scheme="RK2",
                       tune_cache=str(tmp_path / "t.json"))
    S0 = [ens.get_state(i) for i in range(2)]
    ens.step(8)
    for i, r in enumerate(runs):
        orc = NS2D(Oper2D(32, 32), nu_2=r["nu_2"], deltat0=ens.runs[i]["deltat"],
                   type_time_scheme="RK2")
        orc.state_spect[0] = S0[i]
        orc.run(8)
        assert rel_l2(ens.get_state(i), orc.state_spect[0]) <= 1e-12
    ens.close()
