"""GPU parity of the Euler, RK2-trapezoid and phase-shifted schemes
(fluidsim/base/time_stepping/pseudo_spect.py:245-796) against the FP64
oracle.

The device builds each scheme from full tendencies and IF linear
combinations (fsx_if_combine).  For the random variants the translations
the product draws are recorded and handed to the oracle, so both integrate
with the same shifts.  Bars: FP64 state within 1e-11 relative L2 of the
oracle after the steps (CFL on, so dt follows the state); FP32 within 1e-4.
"""

import numpy as np
import pytest

from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3D, NS3DStrat
from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu

SCHEMES = ["Euler", "Euler_phaseshift", "Euler_phaseshift_random", "RK2_trapezoid",
           "RK2_phaseshift", "RK2_phaseshift_random", "RK2_phaseshift_exact"]


def _common(params, scheme, it_end, precision="float64"):
    params.oper.Lx = params.oper.Ly = 2 * np.pi
    if hasattr(params.oper, "Lz"):
        params.oper.Lz = 2 * np.pi
    params.oper.precision = precision
    params.time_stepping.type_time_scheme = scheme
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = it_end
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0


def record_pairs(sim):
    """Wrap the product's pair draw; returns the list it fills."""
    ts = sim.time_stepping
    pairs = []
    if not hasattr(ts, "_get_phaseshift_random"):
        return pairs
    draw = ts._get_phaseshift_random

    def wrapped():
        pair = draw()
        pairs.append(pair)
        return pair

    ts._get_phaseshift_random = wrapped
    return pairs


def replay(pairs):
    it = iter(pairs)
    return lambda: next(it)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_ns2d_scheme_vs_oracle(scheme):
    from fluidsim_amd.solvers.ns2d.solver import Simul

    n, nsteps = 64, 12
    print(f"case ns2d {scheme} n={n} steps={nsteps}", flush=True)
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = n
    params.nu_2 = 1e-3
    _common(params, scheme, nsteps)
    rot0 = ns2d_noise_rot_fft(n, n, seed=21, kmax=12)
    sim = Simul(params)
    sim.state.init_statespect_from(rot_fft=rot0)
    pairs = record_pairs(sim)
    sim.time_stepping.start()
    got = sim.state.get_var_numpy("rot_fft")

    orc = NS2D(Oper2D(n, n), nu_2=1e-3, USE_CFL=True, cfl_coef=0.5,
               type_time_scheme=scheme)
    orc.init_from_rotfft(rot0)
    orc.phase_pairs = replay(pairs)
    orc.run(nsteps)
    if scheme.endswith("_random"):
        assert len(pairs) == nsteps
    err = rel_l2(got, orc.state_spect[0])
    print(f"ns2d {scheme}: rel L2 = {err:.3e}")
    assert sim.time_stepping.t == pytest.approx(orc.t, rel=1e-13)
    assert err <= 1e-11


@pytest.mark.parametrize("scheme", SCHEMES)
def test_ns3d_scheme_vs_oracle(scheme):
    from fluidsim_amd.solvers.ns3d.solver import Simul

    n, nsteps = 32, 5
    print(f"case ns3d {scheme} n={n} steps={nsteps}", flush=True)
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.nu_2 = 1e-2
    _common(params, scheme, nsteps)
    vv = ns3d_spectrum_vv_fft(n, seed=22)
    sim = Simul(params)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    pairs = record_pairs(sim)
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()

    orc = NS3D(Oper3D(n, n, n), nu_2=1e-2, USE_CFL=True, cfl_coef=0.5,
               type_time_scheme=scheme)
    orc.state_spect[:] = np.array(vv)
    orc.phase_pairs = replay(pairs)
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect)
    print(f"ns3d {scheme}: rel L2 = {err:.3e}")
    assert err <= 1e-11


@pytest.mark.parametrize("scheme", ["RK2_phaseshift", "RK2_phaseshift_random"])
def test_strat_scheme_vs_oracle(scheme):
    from fluidsim_amd.solvers.ns3d.strat.solver import Simul

    n, nsteps = 32, 5
    print(f"case strat {scheme} n={n} steps={nsteps}", flush=True)
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.N = 3.0
    params.nu_2 = 1e-2
    _common(params, scheme, nsteps)
    vv = ns3d_spectrum_vv_fft(n, seed=23)
    b = ns3d_spectrum_vv_fft(n, seed=24)[0] * 0.5
    sim = Simul(params)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2], b_fft=b)
    pairs = record_pairs(sim)
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()

    orc = NS3DStrat(Oper3D(n, n, n), N=3.0, nu_2=1e-2, USE_CFL=True, cfl_coef=0.5,
                    type_time_scheme=scheme)
    orc.state_spect[:] = np.array(list(vv) + [b])
    orc.phase_pairs = replay(pairs)
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect)
    print(f"strat {scheme}: rel L2 = {err:.3e}")
    assert err <= 1e-11


def test_random_pairs_properties():
    """get_phases_random (operators2d.py:618-631): per axis the two
    translations differ by half a grid step and alpha lies in
    [-dx/2, dx/2); with nb_pairs = 1 and a new pair every 2 steps the
    second step uses the flipped pair."""
    from fluidsim_amd.solvers.ns3d.solver import Simul

    n = 16
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    _common(params, "RK2_phaseshift_random", 1)
    sim = Simul(params)
    ts = sim.time_stepping
    o = sim.oper
    deltas = (o.deltax, o.deltay, o.deltaz)
    p1 = ts._get_phaseshift_random()
    p2 = ts._get_phaseshift_random()
    assert p2 == (p1[1], p1[0])
    p3 = ts._get_phaseshift_random()
    assert p3 != p1
    for sa, sb in (p1, p3):
        for a, b, d in zip(sa, sb, deltas):
            assert -d / 2 <= a < d / 2
            assert abs(abs(a - b) - d / 2) < 1e-15


def test_ns2d_rk2_phaseshift_fp32():
    from fluidsim_amd.solvers.ns2d.solver import Simul

    n, nsteps = 128, 30
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = n
    params.nu_2 = 1e-3
    _common(params, "RK2_phaseshift", nsteps, precision="float32")
    params.time_stepping.USE_CFL = False
    params.time_stepping.deltat0 = 0.005
    rot0 = ns2d_noise_rot_fft(n, n, seed=25, kmax=16)
    sim = Simul(params)
    sim.state.init_statespect_from(rot_fft=rot0)
    sim.time_stepping.start()
    got = sim.state.get_var_numpy("rot_fft")
    orc = NS2D(Oper2D(n, n), nu_2=1e-3, deltat0=0.005, type_time_scheme="RK2_phaseshift")
    orc.init_from_rotfft(rot0)
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect[0])
    print(f"ns2d RK2_phaseshift FP32 {n}^2 {nsteps} steps: rel L2 = {err:.3e}")
    assert err < 1e-4
