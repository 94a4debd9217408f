"""GPU parity of the forced solvers against the FP64 oracle
(fluidsim/base/forcing/specific.py, anisotropic.py).

The engine and the oracle draw their random seeds and coarse arrays from
numpy's legacy global generator in the same order, so with the same
initial seed they force the same modes with the same raw arrays; the
per-step normalisation runs on the device.  Bars: FP64 state after the
steps within 1e-11 relative L2 of the oracle; injection pin (first-order
injection = forcing_rate) to 1e-10 on the device forcing.
"""

import numpy as np
import pytest

from oracle.forcing import TCRandomForcing
from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3D, NS3DStrat
from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def _common(params, it_end, precision="float64"):
    # the oracle grids are (2 pi)^d; fluidsim's 2D default box is 8 x 8
    params.oper.Lx = params.oper.Ly = 2 * np.pi
    if hasattr(params.oper, "Lz"):
        params.oper.Lz = 2 * np.pi
    params.oper.precision = precision
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = it_end
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0


def injection(oper, v, f, dt):
    return oper.sum_wavenumbers((v.conj() * f).real) + dt / 2 * oper.sum_wavenumbers(
        np.abs(f) ** 2
    )


@pytest.mark.parametrize("fused", [True, False])
def test_ns2d_tcrandom_vs_oracle(fused):
    from fluidsim_amd.solvers.ns2d.solver import Simul

    n, nsteps, seed = 64, 30, 100
    print(f"case ns2d tcrandom n={n} steps={nsteps} fused={fused}", flush=True)
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = n
    params.nu_2 = 1e-3
    _common(params, nsteps)
    params.forcing.enable = True
    params.forcing.type = "tcrandom"
    params.forcing.nkmin_forcing = 4
    params.forcing.nkmax_forcing = 5
    params.forcing.forcing_rate = 1.0
    params.forcing.tcrandom.time_correlation = 0.02
    rot0 = ns2d_noise_rot_fft(n, n, seed=7, kmax=12)
    np.random.seed(seed)
    sim = Simul(params)
    sim.state.init_statespect_from(rot_fft=rot0)
    sim.time_stepping.fused = fused
    sim.time_stepping.start()
    got = sim.state.get_var_numpy("rot_fft")
    F_dev = sim.forcing.get_forcing().to_numpy()

    np.random.seed(seed)
    orc = NS2D(Oper2D(n, n), nu_2=1e-3, USE_CFL=True, cfl_coef=0.5)
    orc.init_from_rotfft(rot0)
    orc.forcing = TCRandomForcing(orc.oper, ("rot_fft",), 4, 5, forcing_rate=1.0,
                                  time_correlation=0.02)
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect[0])
    errF = rel_l2(F_dev, orc._F)
    print(f"ns2d tcrandom: state rel L2 = {err:.3e}, forcing rel L2 = {errF:.3e}")
    assert sim.time_stepping.t == pytest.approx(orc.t, rel=1e-13)
    assert orc.forcing.t_last_change > 0  # the random arrays were renewed
    assert err <= 1e-11 and errF <= 1e-11


def test_ns2d_injection_pin_on_device():
    """One step: the device forcing injects forcing_rate at first order
    with the state at the start of the step and the step's dt."""
    from fluidsim_amd.solvers.ns2d.solver import Simul

    n = 64
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = n
    _common(params, 1)
    params.forcing.enable = True
    params.forcing.type = "tcrandom"
    params.forcing.forcing_rate = 0.37
    np.random.seed(3)
    sim = Simul(params)
    sim.state.init_statespect_from(rot_fft=ns2d_noise_rot_fft(n, n, seed=8, kmax=12))
    rot_before = sim.state.get_var_numpy("rot_fft")
    sim.time_stepping.start()
    dt = sim.time_stepping.deltat
    F = sim.forcing.get_forcing().to_numpy()[0]
    P = injection(Oper2D(n, n), rot_before, F, dt)
    print(f"device injection = {P!r}")
    assert P == pytest.approx(0.37, rel=1e-10)


@pytest.mark.parametrize("fused", [True, False])
def test_ns3d_tcrandom_vx_vs_oracle(fused):
    from fluidsim_amd.solvers.ns3d.solver import Simul

    n, nsteps, seed = 32, 8, 200
    print(f"case ns3d tcrandom vx n={n} fused={fused}", flush=True)
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.nu_2 = 1e-2
    _common(params, nsteps)
    params.forcing.enable = True
    params.forcing.type = "tcrandom"
    params.forcing.key_forced = "vx_fft"
    params.forcing.nkmin_forcing = 2
    params.forcing.nkmax_forcing = 4
    params.forcing.tcrandom.time_correlation = 0.05
    vv = ns3d_spectrum_vv_fft(n, seed=4)
    np.random.seed(seed)
    sim = Simul(params)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    sim.time_stepping.fused = fused
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()

    np.random.seed(seed)
    orc = NS3D(Oper3D(n, n, n), nu_2=1e-2, USE_CFL=True, cfl_coef=0.5)
    orc.state_spect[:] = np.array(vv)
    orc.forcing = TCRandomForcing(orc.oper, NS3D.keys_state_spect, 2, 4,
                                  key_forced="vx_fft", time_correlation=0.05)
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect)
    print(f"ns3d tcrandom vx: rel L2 = {err:.3e}")
    assert err <= 1e-11


def test_strat_tcrandom_anisotropic_vt_vs_oracle():
    """cfg-5-like forcing (toroidal, anisotropic region) at 32^3 FP64."""
    from fluidsim_amd.solvers.ns3d.strat.solver import Simul

    n, nsteps, seed = 32, 6, 300
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.N = 4.0
    params.nu_2 = 1e-2
    _common(params, nsteps)
    params.forcing.enable = True
    params.forcing.type = "tcrandom_anisotropic"
    params.forcing.key_forced = "vt_fft"
    params.forcing.nkmin_forcing = 2
    params.forcing.nkmax_forcing = 4
    params.forcing.tcrandom.time_correlation = 0.05
    vv = ns3d_spectrum_vv_fft(n, seed=5)
    b = ns3d_spectrum_vv_fft(n, seed=6)[0] * 0.5
    np.random.seed(seed)
    sim = Simul(params)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2], b_fft=b)
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()
    assert sim.forcing.forcing_maker.oper_coarse.shapeK == (8, 8, 5)

    np.random.seed(seed)
    orc = NS3DStrat(Oper3D(n, n, n), N=4.0, nu_2=1e-2, USE_CFL=True, cfl_coef=0.5)
    orc.state_spect[:] = np.array(list(vv) + [b])
    orc.forcing = TCRandomForcing(
        orc.oper, NS3DStrat.keys_state_spect, 2, 4, key_forced="vt_fft",
        time_correlation=0.05,
        anisotropic={"angle": np.pi / 4, "delta_angle": np.radians(10)})
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect)
    print(f"strat tcrandom_anisotropic vt: rel L2 = {err:.3e}")
    assert err <= 1e-11


def test_ns2d_forced_fp32_error():
    from fluidsim_amd.solvers.ns2d.solver import Simul

    n, nsteps, seed = 128, 40, 400
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = n
    params.nu_2 = 1e-3
    _common(params, nsteps, precision="float32")
    params.time_stepping.USE_CFL = False
    params.time_stepping.deltat0 = 0.005
    params.forcing.enable = True
    params.forcing.type = "tcrandom"
    params.forcing.nkmin_forcing = 8
    params.forcing.nkmax_forcing = 10
    rot0 = ns2d_noise_rot_fft(n, n, seed=9, kmax=16)
    np.random.seed(seed)
    sim = Simul(params)
    sim.state.init_statespect_from(rot_fft=rot0)
    sim.time_stepping.start()
    got = sim.state.get_var_numpy("rot_fft")
    np.random.seed(seed)
    orc = NS2D(Oper2D(n, n), nu_2=1e-3, deltat0=0.005)
    orc.init_from_rotfft(rot0)
    orc.forcing = TCRandomForcing(orc.oper, ("rot_fft",), 8, 10)
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect[0])
    print(f"ns2d tcrandom FP32 {n}^2 {nsteps} steps: rel L2 = {err:.3e}")
    assert err < 1e-4


def test_in_script_coarse_constant_forcing():
    """A user-set coarse forcing is added unchanged at every stage."""
    from fluidsim_amd.solvers.ns3d.solver import Simul

    n, nsteps = 32, 4
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.nu_2 = 1e-2
    _common(params, nsteps)
    params.time_stepping.USE_CFL = False
    params.time_stepping.deltat0 = 0.01
    params.forcing.enable = True
    params.forcing.type = "in_script_coarse"
    params.forcing.nkmin_forcing = 1
    params.forcing.nkmax_forcing = 3
    vv = ns3d_spectrum_vv_fft(n, seed=10)
    sim = Simul(params)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    maker = sim.forcing.forcing_maker
    oc = Oper3D(*maker.oper_coarse.sizes)
    rng = np.random.default_rng(1)
    fc = np.array([oc.fft(rng.standard_normal(oc.shapeX)) for _ in range(3)])
    for a in fc:
        a[oc.shapeK[0] // 2] = 0
        a[:, oc.shapeK[1] // 2] = 0
        a[..., -1] = 0
    maker.set_forcingc(fc)
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()
    F_full = sim.forcing.get_forcing().to_numpy()

    class Constant:
        def compute(self, state_spect, t, dt):
            return F_full

    orc = NS3D(Oper3D(n, n, n), nu_2=1e-2, deltat0=0.01)
    orc.state_spect[:] = np.array(vv)
    orc.forcing = Constant()
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect)
    print(f"ns3d in_script_coarse: rel L2 = {err:.3e}")
    assert err <= 1e-11
