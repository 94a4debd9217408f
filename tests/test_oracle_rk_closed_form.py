"""Closed-form pins of the oracle's integrating-factor Runge-Kutta steps
(fluidsim/base/time_stepping/pseudo_spect.py:469-517 RK2, 798-984 RK4).

With the affine tendency N(S) = lam S + c (lam, c fixed per mode) one step
of the stage algebra is exactly, with z = lam dt, E = exp(-sigma dt),
E2 = exp(-sigma dt / 2):

  RK4: S' = E (1 + z + z^2/2 + z^3/6 + z^4/24) S
            + dt c [E (1 + z + z^2/2 + z^3/4) / 6 + E2 (2/3 + z/3 + z^2/12) + 1/6]
  RK2: S' = E (1 + z + z^2/2) S + dt c (E z / 2 + E2)

(derived by hand from the stage formulas: the S part is the Taylor
polynomial every 4- / 2-stage method of that order gives; the source part
is specific to the Lawson arrangement of E and E2 in fluidsim's stages, e.g.
Simpson's rule E/6 + 2 E2/3 + 1/6 at z = 0 where an exponential
time-differencing scheme would give (1 - E) / (sigma dt)).  Bars: 1e-13
relative, per mode; a misplaced E2 in one stage (planted in a copy of the
oracle) breaks the pin.
"""

import numpy as np
import pytest

from oracle.operators import Oper2D
from oracle.solvers import NS2D


def _setup(seed=0):
    oper = Oper2D(16, 12, Lx=3.0, Ly=2.0, coef_dealiasing=False)
    orc = NS2D(oper, nu_2=0.05, nu_4=2e-3, nu_m4=0.3)
    rng = np.random.default_rng(seed)
    shape = oper.shapeK
    lam = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 2.0
    c = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    S0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    orc.tendencies_nonlin = lambda S: lam * S + c
    orc.state_spect[0] = S0
    return orc, lam, c, S0


def _closed_form(scheme, orc, lam, c, S0, dt):
    sigma = orc.freq_lin
    E, E2 = np.exp(-dt * sigma), np.exp(-dt / 2 * sigma)
    z = lam * dt
    if scheme == "RK4":
        amp = E * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24)
        src = dt * (E * (1 + z + z**2 / 2 + z**3 / 4) / 6 + E2 * (2 / 3 + z / 3 + z**2 / 12)
                    + 1 / 6)
    else:
        amp = E * (1 + z + z**2 / 2)
        src = dt * (E * z / 2 + E2)
    return amp * S0 + src * c


@pytest.mark.parametrize("scheme", ["RK4", "RK2"])
@pytest.mark.parametrize("dt", [0.01, 0.07, 0.2])
def test_if_rk_step_closed_form(scheme, dt):
    orc, lam, c, S0 = _setup()
    orc.deltat = dt
    getattr(orc, "_time_step_" + scheme)()
    ref = _closed_form(scheme, orc, lam, c, S0, dt)
    err = np.abs(orc.state_spect[0] - ref) / np.maximum(np.abs(ref), 1e-300)
    assert err.max() < 1e-13, err.max()


def test_hypo_viscosity_k0_copy():
    """sigma at k = 0 with nu_m4 copies the first kx mode
    (fluidsim/base/solvers/pseudo_spect.py:176-189), so the mean mode decays
    like kx = dkx, not like 1/1e-28."""
    orc, *_ = _setup()
    f_d, f_d_hypo = orc.compute_freq_diss()
    k2 = orc.oper.deltakx**2
    assert f_d_hypo[0, 1] == pytest.approx(0.3 / k2**2, rel=1e-14)
    assert f_d_hypo[0, 0] == f_d_hypo[0, 1]
    assert f_d[0, 0] == 0.0
    assert np.isfinite(orc.freq_lin).all()


class _MisplacedE2(NS2D):
    """The oracle's RK4 with one planted mistake: the third stage's state
    t1 = S E2 + dt/2 N1 also multiplies N1 by E2."""

    def _time_step_RK4(self):
        dt = self.deltat
        diss, diss2 = self.exact_linear_coefs(dt)
        S = self.state_spect
        N0 = self.tendencies_nonlin(S)
        tmp = (S + dt / 6 * N0) * diss
        t1 = (S + dt / 2 * N0) * diss2
        N1 = self.tendencies_nonlin(t1)
        tmp += dt / 3 * diss2 * N1
        t1 = S * diss2 + dt / 2 * diss2 * N1
        N2 = self.tendencies_nonlin(t1)
        tmp += dt / 3 * diss2 * N2
        t1 = S * diss + dt * diss2 * N2
        N3 = self.tendencies_nonlin(t1)
        S[...] = tmp + dt / 6 * N3


def test_planted_mistake_breaks_the_pin():
    orc, lam, c, S0 = _setup()
    bad = _MisplacedE2(orc.oper, nu_2=0.05, nu_4=2e-3, nu_m4=0.3)
    bad.tendencies_nonlin = orc.tendencies_nonlin
    bad.state_spect[0] = S0
    bad.deltat = 0.07
    bad._time_step_RK4()
    ref = _closed_form("RK4", bad, lam, c, S0, 0.07)
    err = np.abs(bad.state_spect[0] - ref) / np.abs(ref)
    assert err.max() > 1e-6
