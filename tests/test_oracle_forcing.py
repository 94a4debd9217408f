"""Pins of the forcing oracle (oracle/forcing.py) by closed forms.

* the 2nd-degree normalisation makes the first-order injection
  sum_w Re(conj(v) f) + dt/2 sum_w |f|^2 equal forcing_rate exactly
  (specific.py:587-677 docstring);
* with no other term, one explicit step injects exactly forcing_rate dt of
  the quadratic quantity; with the (enstrophy-conserving) ns2d nonlinear
  term the RK4 step injects forcing_rate dt up to O(dt^2);
* forced modes lie in the ring / anisotropic region, the coarse forcing is
  the spectrum of a real field, f(t_last) = f0 and f(t_last + T) = f1;
* the root choice matches numpy.roots (the reference's call);
* a planted mistake (dt/2 dropped from a) breaks the injection pin.
"""

import numpy as np
import pytest

from oracle import forcing as F
from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3DStrat


def injection(oper, v, f, dt):
    return oper.sum_wavenumbers((v.conj() * f).real) + dt / 2 * oper.sum_wavenumbers(
        np.abs(f) ** 2
    )


def random_state(oper, nvar, seed):
    rng = np.random.default_rng(seed)
    return np.array([oper.fft(rng.standard_normal(oper.shapeX)) for _ in range(nvar)])


@pytest.mark.parametrize("rate", [1.0, 0.3])
def test_injection_rate_2d(rate):
    np.random.seed(5)
    o = Oper2D(64, 64)
    forc = F.TCRandomForcing(o, ("rot_fft",), 4, 5, forcing_rate=rate)
    S = random_state(o, 1, 0)
    for t, dt in ((0.0, 0.01), (0.13, 0.003)):
        f = forc.compute(S, t, dt)
        assert injection(o, S[0], f[0], dt) == pytest.approx(rate, rel=1e-12)


def test_injection_rate_3d_components_and_vt():
    o = Oper3D(32, 32, 32)
    S = random_state(o, 4, 1)
    np.random.seed(6)
    fx = F.TCRandomForcing(o, ("vx_fft", "vy_fft", "vz_fft", "b_fft"), 2, 4,
                           key_forced="vx_fft")
    f = fx.compute(S, 0.0, 0.02)
    assert injection(o, S[0], f[0], 0.02) == pytest.approx(1.0, rel=1e-12)
    assert np.all(f[1:] == 0)
    np.random.seed(7)
    fvt = F.TCRandomForcing(o, ("vx_fft", "vy_fft", "vz_fft", "b_fft"), 2, 4,
                            key_forced="vt_fft",
                            anisotropic={"angle": np.pi / 4,
                                         "delta_angle": np.radians(10)})
    f = fvt.compute(S, 0.0, 0.02)
    vt = o.vtfft_from_vecfft(*S[:3])
    ft = o.vtfft_from_vecfft(*f[:3])
    assert injection(o, vt, ft, 0.02) == pytest.approx(1.0, rel=1e-12)
    # the velocity forcing is purely toroidal: no vz, divergence-free
    assert np.all(f[2] == 0) and np.all(f[3] == 0)
    div = o.Kx * f[0] + o.Ky * f[1]
    assert np.abs(div).max() < 1e-12 * np.abs(f[:2]).max()


def test_explicit_step_injects_exactly():
    """rot += dt f with the normalised f raises Z by exactly rate * dt."""
    np.random.seed(8)
    o = Oper2D(64, 64)
    forc = F.TCRandomForcing(o, ("rot_fft",), 4, 5, forcing_rate=0.7)
    S = random_state(o, 1, 2)
    dt = 0.05
    f = forc.compute(S, 0.0, dt)
    Z0 = o.sum_wavenumbers(np.abs(S[0]) ** 2 / 2)
    Z1 = o.sum_wavenumbers(np.abs(S[0] + dt * f[0]) ** 2 / 2)
    assert (Z1 - Z0) / dt == pytest.approx(0.7, rel=1e-11)


def test_rk4_step_injects_enstrophy_ns2d():
    """With nu = 0 the ns2d nonlinear term conserves Z: one forced RK4 step
    raises Z by rate dt + O(dt^2)."""
    o = Oper2D(64, 64)
    rate = 0.5
    errs = []
    for dt in (4e-3, 2e-3):
        np.random.seed(9)
        sim = NS2D(o, deltat0=dt)
        rng = np.random.default_rng(3)
        rot = o.fft(rng.standard_normal(o.shapeX)) * 0.05
        sim.init_from_rotfft(rot)
        sim.forcing = F.TCRandomForcing(o, ("rot_fft",), 4, 5, forcing_rate=rate)
        Z0 = sim.compute_enstrophy()
        sim.one_time_step()
        errs.append(abs((sim.compute_enstrophy() - Z0) / dt - rate))
    assert errs[0] < 1e-5 * rate
    assert errs[1] < errs[0]  # shrinks with dt


def test_forced_region_and_real_field():
    np.random.seed(10)
    o = Oper2D(64, 64)
    forc = F.TCRandomForcing(o, ("rot_fft",), 4, 5)
    oc = forc.oper_c
    forced = ~forc.COND_NO_F
    k = oc.K[forced]
    assert k.min() >= forc.kmin_forcing and k.max() <= forc.kmax_forcing
    # brute-force count of the ring on the coarse half spectrum
    n = 0
    for iy in range(oc.shapeK[0]):
        for ix in range(oc.shapeK[1] - 1):
            if iy == oc.shapeK[0] // 2:
                continue
            kk = np.hypot(oc.KX[iy, ix], oc.KY[iy, ix])
            n += forc.kmin_forcing <= kk <= forc.kmax_forcing
    assert n == forced.sum()
    f0 = forc.forcing0
    x = oc.ifft(f0)
    np.testing.assert_allclose(oc.fft(x), f0, atol=1e-14)


def test_time_correlation_endpoints():
    np.random.seed(11)
    o = Oper2D(32, 32)
    forc = F.TCRandomForcing(o, ("rot_fft",), 2, 3, time_correlation=0.5)
    f0, f1 = forc.forcing0.copy(), forc.forcing1.copy()
    np.testing.assert_allclose(forc.forcingc_raw_each_time(0.0), f0, atol=1e-15)
    np.testing.assert_allclose(forc.forcingc_raw_each_time(0.25),
                               0.5 * (f0 + f1), atol=1e-15)
    # at t_last + T the arrays rotate and the forcing is the old f1
    np.testing.assert_allclose(forc.forcingc_raw_each_time(0.5), f1, atol=1e-15)
    np.testing.assert_allclose(forc.forcing0, f1)


def test_roots_match_numpy():
    rng = np.random.default_rng(4)
    for _ in range(500):
        a = rng.uniform(1e-6, 10)
        b = rng.normal() * 10 ** rng.uniform(-3, 3)
        c = -rng.uniform(1e-3, 10)
        r = np.roots([a, b, c])
        for which, ref in (("minabs", r[1] if abs(r[1]) < abs(r[0]) else r[0]),
                           ("positive", r[1] if r[1] > 0 else r[0])):
            got = F.roots_2nd_degree(a, b, c, which)
            assert got == pytest.approx(ref, rel=1e-11)
    assert F.roots_2nd_degree(0.0, 1.0, -1.0) == 0.0


def test_planted_mistake_breaks_injection(monkeypatch):
    np.random.seed(12)
    o = Oper2D(64, 64)
    forc = F.TCRandomForcing(o, ("rot_fft",), 4, 5)
    S = random_state(o, 1, 5)
    dt = 0.05

    def bad_compute(state_spect, t, dt):
        oc = forc.oper_c
        a_fft = forc.coarse_from_full(state_spect[0])
        fa = forc.forcingc_raw_each_time(t)
        a = oc.sum_wavenumbers(np.abs(fa) ** 2)  # dt/2 dropped
        b = oc.sum_wavenumbers((a_fft.conj() * fa).real)
        fa = fa * F.roots_2nd_degree(a, b, -forc.forcing_rate)
        return np.array([forc.full_from_coarse(fa)])

    f = bad_compute(S, 0.0, dt)
    assert abs(injection(o, S[0], f[0], dt) - 1.0) > 1e-3


def test_strat_vt_forcing_injects_kinetic_energy_rate():
    """ns3d.strat forced through vt_fft: the first-order injection of the
    toroidal kinetic energy equals forcing_rate."""
    np.random.seed(13)
    o = Oper3D(32, 32, 32)
    sim = NS3DStrat(o, N=2.0, deltat0=0.01)
    sim.state_spect[:] = random_state(o, 4, 6) * 0.1
    sim.project_state_spect(sim.state_spect)
    sim.forcing = F.TCRandomForcing(
        o, NS3DStrat.keys_state_spect, 2, 4, key_forced="vt_fft",
        anisotropic={"angle": np.pi / 4, "delta_angle": np.radians(10)})
    f = sim.forcing.compute(sim.state_spect, 0.0, 0.01)
    vt = o.vtfft_from_vecfft(*sim.state_spect[:3])
    assert injection(o, vt, o.vtfft_from_vecfft(*f[:3]), 0.01) == pytest.approx(
        1.0, rel=1e-12)
