"""GPU parity of the non-box truncations and of the hypo-viscous stepping.

Truncations (fluidsim/operators/base.py:47-72, fluidsim/operators/
operators2d.py:171-198, fluidsim/operators/operators3d.py:268-288):
spherical, no_multiple_aliases, NO_SHEAR_MODES and NO_KY0 keep a set of
modes that is not a box, so the plan carries the uint8 mask and the last
forward pass zeroes the masked modes (mask mode of fsx_plan_info).
Hypo-viscosity: sigma = nu_2 k^2 + nu_4 k^4 + nu_m4 / k^4 with the k = 0
value copied from the first kx mode (fluidsim/base/solvers/pseudo_spect.py
:176-189).  Bars: FP64 states within 1e-11 of the oracle after CFL steps
(same dt sequence), FP32 within 1e-5.
"""

import numpy as np
import pytest

from fluidsim_amd import _lib
from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3D, NS3DStrat
from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def _mask_mode(sim):
    info = (_lib.ctypes.c_int * 12)()
    _lib.call("fsx_plan_info", sim.oper.plan.handle, info, 12)
    return info[11]


def _run_ns2d(n, nsteps, oper_kw, precision="float64", use_cfl=True, **visc):
    from fluidsim_amd.solvers.ns2d.solver import Simul

    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = n
    params.oper.Lx = params.oper.Ly = 2 * np.pi
    params.oper.precision = precision
    for k, v in oper_kw.items():
        params.oper[k] = v
    for k, v in visc.items():
        params[k] = v
    params.time_stepping.USE_CFL = use_cfl
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.deltat0 = 0.01
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = nsteps
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    sim = Simul(params)
    rot0 = ns2d_noise_rot_fft(n, n, seed=n + 1, kmax=14)
    orc_oper = Oper2D(n, n, **oper_kw)
    orc_oper.dealiasing(rot0)
    sim.state.init_statespect_from(rot_fft=rot0)
    dts = []
    orig = sim.time_stepping.one_time_step_computation

    def record():
        orig()
        dts.append(sim.time_stepping.deltat)

    sim.time_stepping.one_time_step_computation = record
    sim.time_stepping.start()
    orc = NS2D(orc_oper, USE_CFL=use_cfl, cfl_coef=0.5, deltat0=0.01, **visc)
    orc.init_from_rotfft(rot0)
    odts = orc.run(nsteps, record_dt=True)
    return sim, orc, dts, odts


TRUNC_2D = [{"truncation_shape": "spherical"}, {"truncation_shape": "no_multiple_aliases"},
            {"NO_SHEAR_MODES": True}, {"NO_KY0": True},
            {"truncation_shape": "spherical", "NO_KY0": True}]


@pytest.mark.parametrize("kw", TRUNC_2D)
def test_ns2d_mask_truncations(kw):
    n = 64
    print(f"case ns2d truncation {kw}", flush=True)
    sim, orc, dts, odts = _run_ns2d(n, 10, kw, nu_2=1e-3)
    assert _mask_mode(sim) == 1
    got = sim.state.get_var_numpy("rot_fft")
    err = rel_l2(got, orc.state_spect[0])
    print(f"ns2d {kw}: rel L2 = {err:.3e}")
    np.testing.assert_allclose(dts, odts, rtol=1e-12)
    assert err <= 1e-11
    # the masked modes stay zero
    assert np.abs(got[orc.oper.where_dealiased.astype(bool)]).max() == 0.0


def test_ns2d_mask_truncation_fp32():
    sim, orc, _, _ = _run_ns2d(128, 20, {"truncation_shape": "spherical"},
                               precision="float32", use_cfl=False, nu_2=1e-3)
    err = rel_l2(sim.state.get_var_numpy("rot_fft"), orc.state_spect[0])
    print(f"ns2d FP32 spherical 128^2: rel L2 = {err:.3e}")
    assert err < 1e-5


TRUNC_3D = [{"truncation_shape": "spherical"}, {"truncation_shape": "no_multiple_aliases"},
            {"NO_SHEAR_MODES": True}]


@pytest.mark.parametrize("strat", [False, True])
@pytest.mark.parametrize("kw", TRUNC_3D)
def test_ns3d_mask_truncations(kw, strat):
    if strat:
        from fluidsim_amd.solvers.ns3d.strat.solver import Simul
    else:
        from fluidsim_amd.solvers.ns3d.solver import Simul
    n = 32
    print(f"case ns3d truncation {kw} strat={strat}", flush=True)
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    for k, v in kw.items():
        params.oper[k] = v
    params.nu_2 = 1e-2
    if strat:
        params.N = 2.0
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = 5
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    sim = Simul(params)
    assert _mask_mode(sim) == 1
    op = Oper3D(n, n, n, **kw)
    vv = [a.copy() for a in ns3d_spectrum_vv_fft(n, seed=3)]
    for a in vv:
        op.dealiasing(a)
    if strat:
        b = 0.2 * vv[0]
        sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2], b_fft=b)
        orc = NS3DStrat(op, N=2.0, nu_2=1e-2, USE_CFL=True, cfl_coef=0.5)
        orc.state_spect[:] = np.array(list(vv) + [b])
    else:
        sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
        orc = NS3D(op, nu_2=1e-2, USE_CFL=True, cfl_coef=0.5)
        orc.state_spect[:] = np.array(vv)
    sim.time_stepping.start()
    orc.run(5)
    got = sim.state.state_spect.to_numpy()
    err = rel_l2(got, orc.state_spect)
    print(f"ns3d {kw} strat={strat}: rel L2 = {err:.3e}")
    assert err <= 1e-11
    assert np.abs(got[:, op.where_dealiased.astype(bool)]).max() == 0.0


@pytest.mark.parametrize("scheme", ["RK4", "RK2"])
def test_ns2d_hypoviscosity_steps(scheme):
    """nu_m4 = 0.5 with nu_2, 20 fixed steps (RK4, RK2) within 1e-11."""
    from fluidsim_amd.solvers.ns2d.solver import Simul

    n = 64
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = n
    params.oper.Lx = params.oper.Ly = 2 * np.pi
    params.nu_2 = 1e-3
    params.nu_m4 = 0.5
    params.time_stepping.type_time_scheme = scheme
    params.time_stepping.USE_CFL = False
    params.time_stepping.deltat0 = 0.02
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = 20
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    sim = Simul(params)
    rot0 = ns2d_noise_rot_fft(n, n, seed=9, kmax=6)
    sim.state.init_statespect_from(rot_fft=rot0)
    sim.time_stepping.start()
    orc = NS2D(Oper2D(n, n), nu_2=1e-3, nu_m4=0.5, deltat0=0.02, type_time_scheme=scheme)
    orc.init_from_rotfft(rot0)
    orc.run(20)
    err = rel_l2(sim.state.get_var_numpy("rot_fft"), orc.state_spect[0])
    print(f"ns2d hypo {scheme}: rel L2 = {err:.3e}")
    assert err <= 1e-11


def test_ns3d_hypoviscosity_steps():
    """ns3d 32^3 with nu_m4 = 0.2, nu_4, 6 CFL steps within 1e-11."""
    from fluidsim_amd.solvers.ns3d.solver import Simul

    n = 32
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.nu_4 = 1e-4
    params.nu_m4 = 0.2
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = 6
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    sim = Simul(params)
    vv = ns3d_spectrum_vv_fft(n, seed=5)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    sim.time_stepping.start()
    orc = NS3D(Oper3D(n, n, n), nu_4=1e-4, nu_m4=0.2, USE_CFL=True, cfl_coef=0.5)
    orc.state_spect[:] = np.array(vv)
    orc.run(6)
    err = rel_l2(sim.state.state_spect.to_numpy(), orc.state_spect)
    print(f"ns3d hypo: rel L2 = {err:.3e}")
    assert err <= 1e-11
