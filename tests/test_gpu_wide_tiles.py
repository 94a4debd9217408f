"""GPU checks of the 1024-thread strided tiles that FP32 transforms of
length >= 1024 use when a 512-thread tile would hold fewer than 8 lines
(the multi-transform z-passes of 1024^3 runs).

Small grids never take that path (small problems shrink the lines per tile
to fill the chip), so the case is a 256 x 256 x 1024 box: its z-passes of
length 1024 hold 8 lines. The FP32 state after fused RK4 steps (curl z-pass
and Runge-Kutta z-pass in wide tiles) is compared with the FP64 engine
(512-thread tiles, held to the FP64 oracle at 1e-12 by the other tests) at
the FP32 bar of 2e-6 relative L2 per step; ns3d and ns3d.strat.
"""

import numpy as np
import pytest

from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def _sim(Simul, precision, vv, dt, strat):
    params = Simul.create_default_params()
    po = params.oper
    po.nx, po.ny, po.nz = 256, 256, 1024
    po.Lx = po.Ly = po.Lz = 2 * np.pi
    po.precision = precision
    params.nu_2 = 1e-3
    if strat:
        params.N = 4.0
    params.time_stepping.USE_CFL = False
    params.time_stepping.deltat0 = dt
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    sim = Simul(params)
    cdt = np.complex64 if precision == "float32" else np.complex128
    kw = {k: a.astype(cdt) for k, a in zip(("vx_fft", "vy_fft", "vz_fft"), vv)}
    if strat:
        kw["b_fft"] = (0.5 * vv[0]).astype(cdt)
    sim.state.init_statespect_from(**kw)
    return sim


@pytest.mark.parametrize("strat", [False, True])
def test_wide_tiles_fp32_vs_fp64(strat):
    if strat:
        from fluidsim_amd.solvers.ns3d.strat.solver import Simul
    else:
        from fluidsim_amd.solvers.ns3d.solver import Simul
    print(f"case wide tiles 256x256x1024 strat={strat}", flush=True)
    vv = ns3d_spectrum_vv_fft(256, 256, 1024, seed=11)
    nsteps, dt = 2, 2e-3
    out = {}
    for prec in ("float64", "float32"):
        sim = _sim(Simul, prec, vv, dt, strat)
        for _ in range(nsteps):
            sim.time_stepping.one_time_step_computation()
        out[prec] = sim.state.state_spect.to_numpy().astype(np.complex128)
        del sim
    err = rel_l2(out["float32"], out["float64"])
    print(f"FP32 (wide tiles) vs FP64: rel L2 = {err:.2e}")
    assert err < 2e-6 * nsteps
    assert np.abs(out["float64"]).max() > 0
