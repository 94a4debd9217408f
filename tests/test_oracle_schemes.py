"""Pins of the Euler / RK2 / phase-shifted schemes of the oracle
(pseudo_spect.py:245-796).

* with no nonlinear term every scheme reproduces the exact decay
  exp(-sigma t) (integrating factor);
* order of convergence on ns2d: Euler variants first order, RK2 variants
  second order (error ratio at dt and dt/2 against a fine RK4 reference);
* with the 2/3 truncation there is no aliasing, so a phase-shifted
  tendency equals the plain one: RK2_phaseshift reduces to
  S E + dt E2 (N0 + N(S1)) / 2 and RK2_phaseshift_exact to the same with
  N_d0 = N0;
* a planted mistake (shift not undone) breaks the shifted-tendency pin.
"""

import numpy as np
import pytest

from oracle.operators import Oper2D
from oracle.solvers import NS2D
from fluidsim_amd.synthetic import ns2d_noise_rot_fft

SCHEMES = ["Euler", "Euler_phaseshift", "Euler_phaseshift_random", "RK2_trapezoid",
           "RK2", "RK2_phaseshift", "RK2_phaseshift_random", "RK2_phaseshift_exact"]


def fixed_pairs(sim):
    rng = np.random.default_rng(0)

    def pairs():
        a = rng.uniform(-0.5, 0.5, 2)
        b = np.where(a < 0, a + 0.5, a - 0.5)
        o = sim.oper
        return ((a[0] * o.deltax, a[1] * o.deltay, 0.0),
                (b[0] * o.deltax, b[1] * o.deltay, 0.0))

    return pairs


@pytest.mark.parametrize("scheme", SCHEMES)
def test_linear_decay_exact(scheme):
    o = Oper2D(32, 32)
    sim = NS2D(o, nu_2=0.05, deltat0=0.1, type_time_scheme=scheme)
    sim.tendencies_nonlin = lambda S: np.zeros_like(S)
    sim.phase_pairs = fixed_pairs(sim)
    rot = ns2d_noise_rot_fft(32, 32, seed=1, kmax=8)
    sim.init_from_rotfft(rot)
    S0 = sim.state_spect.copy()
    sim.run(5)
    np.testing.assert_allclose(sim.state_spect, S0 * np.exp(-0.05 * o.K2 * 0.5),
                               rtol=1e-12, atol=1e-15)


def error_at(scheme, dt, T=0.08):
    o = Oper2D(32, 32)
    rot = ns2d_noise_rot_fft(32, 32, seed=2, kmax=8)
    sim = NS2D(o, nu_2=1e-3, deltat0=dt, type_time_scheme=scheme)
    sim.phase_pairs = fixed_pairs(sim)
    sim.init_from_rotfft(rot)
    sim.run(int(round(T / dt)))
    ref = NS2D(o, nu_2=1e-3, deltat0=T / 256)
    ref.init_from_rotfft(rot)
    ref.run(256)
    return np.linalg.norm(sim.state_spect - ref.state_spect) / np.linalg.norm(ref.state_spect)


@pytest.mark.parametrize("scheme,order", [("Euler", 1), ("Euler_phaseshift", 1),
                                          ("RK2_trapezoid", 2), ("RK2_phaseshift", 2),
                                          ("RK2_phaseshift_exact", 2)])
def test_order(scheme, order):
    e1, e2 = error_at(scheme, 0.02), error_at(scheme, 0.01)
    ratio = e1 / e2
    assert 2 ** order * 0.8 < ratio < 2 ** order * 1.25, ratio


def test_phaseshift_reduces_without_aliasing():
    o = Oper2D(32, 32)
    rot = ns2d_noise_rot_fft(32, 32, seed=3, kmax=8)
    dt = 0.02
    a = NS2D(o, nu_2=1e-3, deltat0=dt, type_time_scheme="RK2_phaseshift")
    a.init_from_rotfft(rot)
    S = a.state_spect.copy()
    ps = a.phaseshift(a.half_shift())
    N0 = a.tendencies_nonlin(S)
    np.testing.assert_allclose(a.tendencies_nonlin(ps * S) / ps, N0, atol=1e-13)
    a.run(1)
    E, E2 = a.exact_linear_coefs(dt)
    S1 = (S + dt * N0) * E
    manual = S * E + dt * E2 * 0.5 * (N0 + a.tendencies_nonlin(S1))
    np.testing.assert_allclose(a.state_spect, manual, atol=1e-13)
    b = NS2D(o, nu_2=1e-3, deltat0=dt, type_time_scheme="RK2_phaseshift_exact")
    b.init_from_rotfft(rot)
    b.run(1)
    np.testing.assert_allclose(b.state_spect, manual, atol=1e-13)


def test_planted_mistake_shift_not_undone():
    o = Oper2D(32, 32)
    a = NS2D(o, nu_2=1e-3)
    a.init_from_rotfft(ns2d_noise_rot_fft(32, 32, seed=4, kmax=8))
    S = a.state_spect.copy()
    ps = a.phaseshift(a.half_shift())
    N0 = a.tendencies_nonlin(S)
    assert np.abs(a.tendencies_nonlin(ps * S) - N0).max() > 1e-3 * np.abs(N0).max()
