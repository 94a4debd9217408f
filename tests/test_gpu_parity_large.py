"""FP64 parity at larger grids than the unit tests: the engine's state after
N adaptive-CFL RK4 steps against the FP64 oracle (relative L2 <= 1e-11),
with the same dt sequence (the CFL rule with its 2 % hysteresis,
fluidsim/base/time_stepping/base.py:341-354).

* ns3d 128^3, nu_8 with k_max eta = 1, isotropic unit-energy init, 10 steps;
* ns3d.strat 64^3, N = 4, nu_8, velocity and buoyancy init, 8 steps.
"""

import numpy as np
import pytest

from oracle.operators import Oper3D
from oracle.solvers import NS3D, NS3DStrat
from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft, nu8_for_kmax_eta

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def _params(Simul, n, nsteps):
    params = Simul.create_default_params()
    po = params.oper
    po.nx = po.ny = po.nz = n
    po.Lx = po.Ly = po.Lz = 2 * np.pi
    po.precision = "float64"
    params.nu_8 = nu8_for_kmax_eta(n)
    pt = params.time_stepping
    pt.USE_CFL = True
    pt.cfl_coef = 0.5
    pt.USE_T_END = False
    pt.it_end = nsteps
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    return params


def _engine_dts(sim, nsteps):
    ts = sim.time_stepping
    ts.prepare_main_loop()
    dts = []
    for _ in range(nsteps):
        ts.one_time_step()
        dts.append(ts.deltat)
    return dts


def test_ns3d_128_fp64_cfl_steps_vs_oracle():
    from fluidsim_amd.solvers.ns3d.solver import Simul

    n, nsteps = 128, 10
    print(f"case ns3d {n}^3 FP64 {nsteps} CFL steps", flush=True)
    vv = ns3d_spectrum_vv_fft(n, seed=21)
    sim = Simul(_params(Simul, n, nsteps))
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    dts = _engine_dts(sim, nsteps)
    got = sim.state.state_spect.to_numpy()
    orc = NS3D(Oper3D(n, n, n), nu_8=nu8_for_kmax_eta(n), USE_CFL=True, cfl_coef=0.5)
    orc.state_spect[:] = np.array(vv)
    odts = orc.run(nsteps, record_dt=True)
    np.testing.assert_allclose(dts, odts, rtol=1e-12)
    err = rel_l2(got, orc.state_spect)
    print(f"ns3d {n}^3 FP64 rel L2 = {err:.2e}")
    assert err <= 1e-11


def test_strat_64_fp64_cfl_steps_vs_oracle():
    from fluidsim_amd.solvers.ns3d.strat.solver import Simul

    n, nsteps = 64, 8
    print(f"case strat {n}^3 FP64 {nsteps} CFL steps", flush=True)
    vv = ns3d_spectrum_vv_fft(n, seed=22)
    b = 0.5 * ns3d_spectrum_vv_fft(n, seed=23)[0]
    params = _params(Simul, n, nsteps)
    params.N = 4.0
    sim = Simul(params)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2], b_fft=b)
    dts = _engine_dts(sim, nsteps)
    got = sim.state.state_spect.to_numpy()
    orc = NS3DStrat(Oper3D(n, n, n), N=4.0, nu_8=nu8_for_kmax_eta(n), USE_CFL=True,
                    cfl_coef=0.5)
    orc.state_spect[:] = np.array(list(vv) + [b])
    odts = orc.run(nsteps, record_dt=True)
    np.testing.assert_allclose(dts, odts, rtol=1e-12)
    err = rel_l2(got, orc.state_spect)
    print(f"strat {n}^3 FP64 rel L2 = {err:.2e}")
    assert err <= 1e-11
