"""Pins of the oracle's tendencies and time stepping (CPU only)."""

import numpy as np
import pytest

from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3D, NS3DStrat
from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft


def grid2d(oper):
    y, x = np.meshgrid(
        np.arange(oper.ny) * oper.deltay, np.arange(oper.nx) * oper.deltax,
        indexing="ij",
    )
    return x, y


def grid3d(oper):
    z, y, x = np.meshgrid(
        np.arange(oper.nz) * oper.deltaz,
        np.arange(oper.ny) * oper.deltay,
        np.arange(oper.nx) * oper.deltax,
        indexing="ij",
    )
    return x, y, z


def make_ns2d(n=32, **kw):
    oper = Oper2D(n, n)
    sim = NS2D(oper, **kw)
    sim.init_from_rotfft(ns2d_noise_rot_fft(n, n, seed=3, kmax=n / 4))
    return sim


# ---------------------------------------------------------------- ns2d ---
def test_ns2d_tendency_conserves_energy_and_enstrophy():
    """fluidsim/solvers/ns2d/test_solver.py:56-69 and 181-193."""
    sim = make_ns2d(32)
    rot_fft = sim.state_spect[0]
    F = sim.tendencies_nonlin(sim.state_spect)[0]
    sw = sim.oper.sum_wavenumbers
    T_rot = np.real(F.conj() * rot_fft)
    assert abs(sw(T_rot) / sw(np.abs(T_rot))) < 1e-15
    psi = rot_fft / sim.oper.K2_not0
    T_E = np.real(F.conj() * psi)
    assert abs(sw(T_E) / sw(np.abs(T_E))) < 1e-14


def test_ns2d_taylor_green_is_steady():
    """rot = 2 sin x sin y is an exact steady Euler solution: N == 0."""
    oper = Oper2D(16, 16)
    sim = NS2D(oper)
    x, y = grid2d(oper)
    sim.init_from_rotfft(oper.fft(2 * np.sin(x) * np.sin(y)))
    F = sim.tendencies_nonlin(sim.state_spect)[0]
    assert np.abs(F).max() < 1e-15


def test_ns2d_two_mode_closed_form():
    """rot = cos x + 4 cos 2y  =>  u = -2 sin 2y, v = sin x and
    N = -u d_x rot - v d_y rot = 6 sin x sin 2y (pins signs and operands)."""
    oper = Oper2D(16, 16)
    sim = NS2D(oper)
    x, y = grid2d(oper)
    sim.init_from_rotfft(oper.fft(np.cos(x) + 4 * np.cos(2 * y)))
    F = sim.tendencies_nonlin(sim.state_spect)[0]
    np.testing.assert_allclose(oper.ifft(F), 6 * np.sin(x) * np.sin(2 * y), atol=1e-13)
    # beta term: -v * beta = -beta sin x
    sim.beta = 0.5
    F = sim.tendencies_nonlin(sim.state_spect)[0]
    np.testing.assert_allclose(
        oper.ifft(F), 6 * np.sin(x) * np.sin(2 * y) - 0.5 * np.sin(x), atol=1e-13
    )


def test_planted_mistakes_fail_the_ns2d_pins(monkeypatch):
    """Each plausible mistake planted in the oracle fails a pin above."""
    oper = Oper2D(16, 16)
    x, y = grid2d(oper)
    rot0 = oper.fft(np.cos(x) + 4 * np.cos(2 * y))
    expected = 6 * np.sin(x) * np.sin(2 * y)

    def run():
        sim = NS2D(oper)
        sim.init_from_rotfft(rot0)
        return oper.ifft(sim.tendencies_nonlin(sim.state_spect)[0])

    orig = Oper2D.vecfft_from_rotfft
    # 1) velocity with the wrong sign
    monkeypatch.setattr(
        Oper2D, "vecfft_from_rotfft",
        lambda self, r: tuple(-a for a in orig(self, r)),
    )
    assert not np.allclose(run(), expected, atol=1e-10)
    # 2) transposed velocity components
    monkeypatch.setattr(
        Oper2D, "vecfft_from_rotfft", lambda self, r: orig(self, r)[::-1]
    )
    assert not np.allclose(run(), expected, atol=1e-10)
    monkeypatch.setattr(Oper2D, "vecfft_from_rotfft", orig)
    # 3) missing dealiasing breaks exact enstrophy conservation
    sim = make_ns2d(32)
    monkeypatch.setattr(Oper2D, "dealiasing", lambda self, *a: None)
    rng = np.random.default_rng(0)
    rot = sim.oper.fft(rng.standard_normal(sim.oper.shapeX))
    rot[0, 0] = 0
    sim.state_spect[0] = rot
    F = sim.tendencies_nonlin(sim.state_spect)[0]
    T = np.real(F.conj() * rot)
    sw = sim.oper.sum_wavenumbers
    assert abs(sw(T) / sw(np.abs(T))) > 1e-12


# ---------------------------------------------------------------- ns3d ---
def test_ns3d_tendency_conserves_energy():
    """fluidsim/solvers/ns3d/test_solver.py:80-96."""
    oper = Oper3D(16, 12, 8, 2 * np.pi, 3.0, 4.0)
    sim = NS3D(oper)
    rng = np.random.default_rng(1)
    v = [oper.fft(rng.standard_normal(oper.shapeX)) for _ in range(3)]
    oper.project_perpk3d(*v)
    oper.dealiasing(*v)
    sim.state_spect[:] = np.array(v)
    F = sim.tendencies_nonlin(sim.state_spect)
    T = sum(np.real(F[i].conj() * sim.state_spect[i]) for i in range(3))
    sw = oper.sum_wavenumbers
    assert abs(sw(T) / sw(np.abs(T))) < 1e-15


def test_ns3d_abc_flow_is_steady():
    """ABC flow: omega = v, so v x omega = 0 and N == 0."""
    oper = Oper3D(16, 16, 16)
    sim = NS3D(oper)
    x, y, z = grid3d(oper)
    A, B, C = 1.0, 0.7, 0.3
    v = (
        A * np.sin(z) + C * np.cos(y),
        B * np.sin(x) + A * np.cos(z),
        C * np.sin(y) + B * np.cos(x),
    )
    sim.state_spect[:] = np.array([oper.fft(a) for a in v])
    F = sim.tendencies_nonlin(sim.state_spect)
    assert np.abs(F).max() < 1e-15


def test_ns3d_shear_closed_form():
    """v = (cos y, 0, cos x): -(v.grad)v = (0, 0, sin x cos y) and the rest
    is a gradient, so N = (0, 0, sin x cos y) (pins v x omega vs omega x v,
    the curl sign and the projection)."""
    oper = Oper3D(16, 16, 16)
    sim = NS3D(oper)
    x, y, z = grid3d(oper)
    v = (np.cos(y), np.zeros_like(x), np.cos(x))
    sim.state_spect[:] = np.array([oper.fft(a) for a in v])
    F = sim.tendencies_nonlin(sim.state_spect)
    np.testing.assert_allclose(oper.ifft(F[0]), 0, atol=1e-14)
    np.testing.assert_allclose(oper.ifft(F[1]), 0, atol=1e-14)
    np.testing.assert_allclose(oper.ifft(F[2]), np.sin(x) * np.cos(y), atol=1e-14)
    # Coriolis: omega_z += f at k = 0 adds v x (0,0,f) = f (vy, -vx, 0).
    # For v = (cos z, 0, 0) the advection term is a gradient and
    # f (vy, -vx, 0) = (0, -f cos z, 0) is solenoidal, so N = (0, -f cos z, 0).
    sim.f = 2.0
    v = (np.cos(z), np.zeros_like(x), np.zeros_like(x))
    sim.state_spect[:] = np.array([oper.fft(a) for a in v])
    F = sim.tendencies_nonlin(sim.state_spect)
    np.testing.assert_allclose(oper.ifft(F[0]), 0, atol=1e-14)
    np.testing.assert_allclose(oper.ifft(F[1]), -2 * np.cos(z), atol=1e-13)
    np.testing.assert_allclose(oper.ifft(F[2]), 0, atol=1e-14)


def test_strat_energy_and_linear_wave():
    """Energy incl. E_A = b^2/(2N^2) is conserved by the tendency
    (fluidsim/solvers/ns3d/strat/test_solver.py:35-58); a kz = 0 mode
    (vz, b) = (a cos x, beta cos x) obeys d_t vz = b, d_t b = -N^2 vz."""
    oper = Oper3D(12, 12, 8)
    N = 2.0
    sim = NS3DStrat(oper, N=N)
    rng = np.random.default_rng(2)
    v = [oper.fft(rng.standard_normal(oper.shapeX)) for _ in range(4)]
    oper.project_perpk3d(*v[:3])
    oper.dealiasing(*v)
    sim.state_spect[:] = np.array(v)
    F = sim.tendencies_nonlin(sim.state_spect)
    S = sim.state_spect
    T = sum(np.real(F[i].conj() * S[i]) for i in range(3))
    T = T + np.real(F[3].conj() * S[3]) / N**2
    sw = oper.sum_wavenumbers
    assert abs(sw(T) / sw(np.abs(T))) < 1e-14
    x, y, z = grid3d(oper)
    a, beta = 0.3, 0.7
    fields = [np.zeros_like(x), np.zeros_like(x), a * np.cos(x), beta * np.cos(x)]
    sim.state_spect[:] = np.array([oper.fft(f) for f in fields])
    F = sim.tendencies_nonlin(sim.state_spect)
    np.testing.assert_allclose(oper.ifft(F[0]), 0, atol=1e-14)
    np.testing.assert_allclose(oper.ifft(F[1]), 0, atol=1e-14)
    np.testing.assert_allclose(oper.ifft(F[2]), beta * np.cos(x), atol=1e-14)
    np.testing.assert_allclose(oper.ifft(F[3]), -(N**2) * a * np.cos(x), atol=1e-13)


# ------------------------------------------------------ time stepping ---
def test_rk4_exact_decay_without_nonlinearity():
    """Taylor-Green (N == 0): the integrating factor makes the scheme exact,
    rot(t) = rot0 exp(-sigma t) with sigma = nu2 k^2 + nu4 k^4 + nu8 k^8,
    here k^2 = 2, for any dt (pins sigma and the exponent sign)."""
    oper = Oper2D(16, 16)
    nu2, nu4, nu8 = 0.05, 0.01, 0.001
    for scheme in ("RK4", "RK2"):
        sim = NS2D(oper, nu_2=nu2, nu_4=nu4, nu_8=nu8, type_time_scheme=scheme,
                   deltat0=0.37)
        x, y = grid2d(oper)
        rot0 = 2 * np.sin(x) * np.sin(y)
        sim.init_from_rotfft(oper.fft(rot0))
        sim.run(5)
        sigma = nu2 * 2 + nu4 * 4 + nu8 * 16
        np.testing.assert_allclose(
            oper.ifft(sim.state_spect[0]), rot0 * np.exp(-sigma * sim.t), atol=1e-14
        )


def _ns2d_error_vs_dt(scheme, dts, T=0.4):
    oper = Oper2D(32, 32)
    rot0 = ns2d_noise_rot_fft(32, 32, seed=5, kmax=6)
    res = []
    for dt in dts:
        sim = NS2D(oper, nu_2=1e-2, type_time_scheme=scheme, deltat0=dt)
        sim.init_from_rotfft(rot0)
        sim.run(int(round(T / dt)))
        res.append(sim.state_spect[0].copy())
    ref = res[-1]
    return [np.linalg.norm(r - ref) for r in res[:-1]]


def test_rk4_and_rk2_convergence_order():
    e = _ns2d_error_vs_dt("RK4", [0.1, 0.05, 0.025, 0.0015625])
    assert 12 < e[0] / e[1] < 20 and 12 < e[1] / e[2] < 20
    e = _ns2d_error_vs_dt("RK2", [0.1, 0.05, 0.025, 0.0015625])
    assert 3 < e[0] / e[1] < 5 and 3 < e[1] / e[2] < 5


def test_rk4_stage_algebra_linear_closed_form():
    """With N(S) = lam S the IF-RK4 step equals the exact exp((lam-s)dt) up to
    O(dt^5) per step; a dropped or misplaced E/E2 factor breaks the order."""
    oper = Oper2D(8, 8)

    class Lin(NS2D):
        lam = -0.8 + 0.5j

        def tendencies_nonlin(self, state_spect):
            return self.lam * state_spect

    errs = []
    for dt in (0.2, 0.1):
        sim = Lin(oper, nu_2=0.3, deltat0=dt)
        sim.state_spect[0] = 1.0
        sim.one_time_step_computation_nodealias = None
        S0 = sim.state_spect.copy()
        sim._time_step_RK4()
        sigma = sim.freq_lin
        exact = S0 * np.exp((Lin.lam - sigma) * dt)
        errs.append(np.abs(sim.state_spect - exact).max())
    assert 24 < errs[0] / errs[1] < 40  # ~2^5


def test_cfl_hysteresis():
    """base.py:341-354: dt = min(CFL/tmp, dt_max), changed only when the
    relative change exceeds 2 %."""
    oper = Oper2D(8, 8)

    class Scripted(NS2D):
        seq = []

        def cfl_tmp(self):
            return self.seq.pop(0)

    sim = Scripted(oper, USE_CFL=True, cfl_coef=0.5, deltat0=0.2, deltat_max=0.2)
    Scripted.seq = [10.0, 10.1, 10.3, 0.0, 1.0]
    sim.compute_time_increment_CLF()
    assert sim.deltat == 0.05
    sim.compute_time_increment_CLF()  # 0.5/10.1: 1 % change -> kept
    assert sim.deltat == 0.05
    sim.compute_time_increment_CLF()  # 0.5/10.3: 2.9 % -> updated
    assert sim.deltat == 0.5 / 10.3
    sim.compute_time_increment_CLF()  # tmp = 0 -> deltat_max
    assert sim.deltat == 0.2
    sim.compute_time_increment_CLF()  # 0.5 > dt_max -> dt_max
    assert sim.deltat == 0.2


def test_energy_budget_closes():
    """Without forcing dE/dt = -eps: the centred difference matches the mean
    of eps at both ends to O(dt^2) (and the scheme error is O(dt^4))."""
    oper = Oper2D(32, 32)
    rot0 = ns2d_noise_rot_fft(32, 32, seed=7, kmax=8)
    errs = []
    for dt in (0.02, 0.01):
        sim = NS2D(oper, nu_2=2e-2, deltat0=dt)
        sim.init_from_rotfft(rot0)
        m0 = sim.spatial_means()
        sim.run(1)
        m1 = sim.spatial_means()
        dEdt = (m1["E"] - m0["E"]) / dt
        eps = 0.5 * (m0["epsK"] + m1["epsK"])
        errs.append(abs(dEdt + eps) / eps)
    assert errs[0] < 1e-3
    assert 3 < errs[0] / errs[1] < 5


def test_ns3d_spectrum_init_is_solenoidal_unit_energy():
    vv = ns3d_spectrum_vv_fft(16, seed=1)
    oper = Oper3D(16, 16, 16)
    kdotv = oper.Kx * vv[0] + oper.Ky * vv[1] + oper.Kz * vv[2]
    assert np.abs(kdotv).max() < 1e-14
    E = oper.sum_wavenumbers(0.5 * sum(np.abs(a) ** 2 for a in vv))
    assert E == pytest.approx(1.0, rel=1e-13)
    # Hermitian: round trip through real space is exact
    for a in vv:
        np.testing.assert_allclose(oper.fft(oper.ifft(a)), a, atol=1e-15)


def test_ns2d_noise_init_velo_max():
    oper = Oper2D(64, 64)
    rot = ns2d_noise_rot_fft(64, 64, seed=0)
    sim = NS2D(oper)
    sim.init_from_rotfft(rot)
    ph = sim.state_phys()
    vmax = np.sqrt(ph["ux"] ** 2 + ph["uy"] ** 2).max()
    # the final dealiasing of rot removes nothing (|k| <= 16 < 2/3 * 32)
    assert vmax == pytest.approx(1.0, rel=1e-12)
