"""Pins of the diagnostics oracle (oracle/diagnostics.py).

* binning: a brute-force loop over the FULL grid (both signs of kx, weight
  1 per mode) reproduces the half-spectrum histograms with their Parseval
  weights (1D, shell, kz-kh; 2D and 3D; odd and even sizes; unequal boxes);
* sum rules: sum_j S_j dk = sum_wavenumbers(q) for every histogram;
* single-mode closed form of the linear share between two shells;
* budgets: under the oracle's own Euler step (nu = 0) the spectra change by
  dt times the transfer terms to first order (error ratio 2 when dt
  halves): ns2d E and Z, ns3d K, Kz and Khr, strat E_K = T_K - C and
  E_A = T_A + C; with the nonlinear term off the dissipation terms give
  the viscous decay;
* spatial means: injection PK1 + PK2 (ns3d) and PZ1 + PZ2 (ns2d) equal the
  forcing rate of the normalised forcing oracle; E equals the sum of the
  spectra;
* planted mistakes (weights dropped, transfer sign flipped) fail the pins.
"""

import itertools

import numpy as np
import pytest

from oracle import diagnostics as D
from oracle.forcing import TCRandomForcing
from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3D, NS3DStrat


def real_spectrum(oper, rng, nvar=1, scale=1.0):
    return np.array([oper.fft(rng.standard_normal(oper.shapeX)) * scale
                     for _ in range(nvar)])


def smooth_state(oper, rng, nvar, kmax):
    """Band-limited, dealiased, Hermitian-consistent random fields."""
    S = real_spectrum(oper, rng, nvar)
    S[:, np.sqrt(oper.K2) > kmax] = 0
    oper.dealiasing(S)
    return np.array([oper.fft(oper.ifft(a)) for a in S])


def ksig(i, n):
    return i if i <= n // 2 else i - n


def brute_force(oper, X):
    """Histograms by a loop over the full grid of fftn(X) / N."""
    b = D.Bins(oper)
    F = np.fft.fftn(X) / X.size
    q = np.abs(F) ** 2 / 2
    dim = X.ndim
    sx, sy, sz = np.zeros(b.nkx), np.zeros(b.nky), np.zeros(b.nkz)
    shell = np.zeros(b.nk + 1)
    kzkh = np.zeros((b.nkz, b.nkh + 1))
    for idx in itertools.product(*(range(n) for n in X.shape)):
        if dim == 2:
            iy, ix = idx
            iz = 0
        else:
            iz, iy, ix = idx
        ikx, iky = ksig(ix, oper.nx), ksig(iy, oper.ny)
        ikz = ksig(iz, oper.nz) if dim == 3 else 0
        val = q[idx]
        kx, ky = ikx * oper.deltakx, iky * oper.deltaky
        kz = ikz * oper.deltakz if dim == 3 else 0.0
        sx[abs(ikx)] += val
        sy[abs(iky)] += val
        sz[abs(ikz)] += val
        kk = np.sqrt(kx**2 + ky**2 + kz**2) if dim == 3 else np.sqrt(kx**2 + ky**2)
        j = int(kk / b.dk)
        if j >= b.nk - 1:
            shell[b.nk - 1] += val
        else:
            c = (kk - j * b.dk) / b.dk
            shell[j] += (1 - c) * val
            shell[j + 1] += c * val
        if dim == 3:
            kh = np.sqrt(kx**2 + ky**2)
            j = int(kh / b.dkh)
            if j >= b.nkh - 1:
                kzkh[abs(ikz), b.nkh - 1] += val
            else:
                c = (kh - j * b.dkh) / b.dkh
                kzkh[abs(ikz), j] += (1 - c) * val
                kzkh[abs(ikz), j + 1] += c * val
    out = {"kx": sx / oper.deltakx, "ky": sy / oper.deltaky, "shell": shell[: b.nk] / b.dk}
    if dim == 3:
        out["kz"] = sz / oper.deltakz
        out["kzkh"] = kzkh[:, : b.nkh] / (oper.deltakz * b.dkh)
    return out


OPERS = [
    lambda: Oper2D(8, 6, Lx=3.0, Ly=5.0),
    lambda: Oper2D(9, 8),
    lambda: Oper3D(8, 6, 4, Lx=2 * np.pi, Ly=3.0, Lz=7.0),
    lambda: Oper3D(7, 8, 6),
    lambda: Oper3D(8, 8, 8, Lx=4 * np.pi),
]


@pytest.mark.parametrize("make", OPERS)
def test_binning_matches_full_grid_brute_force(make):
    oper = make()
    X = np.random.default_rng(0).standard_normal(oper.shapeX)
    q = np.abs(oper.fft(X)) ** 2 / 2
    ref = brute_force(oper, X)
    got = D.spectra_1d(q, oper)
    np.testing.assert_allclose(got[0], ref["kx"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(got[1], ref["ky"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(D.shell_spectrum(q, oper), ref["shell"], rtol=1e-12,
                               atol=1e-15)
    if X.ndim == 3:
        np.testing.assert_allclose(got[2], ref["kz"], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(D.spectrum_kzkh(q, oper), ref["kzkh"], rtol=1e-12,
                                   atol=1e-15)


@pytest.mark.parametrize("make", OPERS)
def test_sum_rules(make):
    oper = make()
    rng = np.random.default_rng(1)
    q = np.abs(real_spectrum(oper, rng)[0]) ** 2
    total = oper.sum_wavenumbers(q)
    b = D.Bins(oper)
    s1 = D.spectra_1d(q, oper)
    dks = (oper.deltakx, oper.deltaky) + ((oper.deltakz,) if q.ndim == 3 else ())
    for s, dk in zip(s1, dks):
        assert s.sum() * dk == pytest.approx(total, rel=1e-13)
    assert D.shell_spectrum(q, oper).sum() * b.dk == pytest.approx(total, rel=1e-13)
    if q.ndim == 3:
        kzkh = D.spectrum_kzkh(q, oper)
        assert kzkh.sum() * oper.deltakz * b.dkh == pytest.approx(total, rel=1e-13)


def test_single_mode_share():
    """One mode at kx = 2.5 (dkx = 1/2, dk = 1): half of its weighted value
    in shells 2 and 3, all of it in kx bin 5."""
    oper = Oper3D(16, 8, 8, Lx=4 * np.pi)
    b = D.Bins(oper)
    assert (b.dk, b.nk, b.dkh, b.nkh) == (1.0, 4, 1.0, 4)
    q = np.zeros(oper.shapeK)
    q[0, 0, 5] = 3.0
    shell = D.shell_spectrum(q, oper)
    np.testing.assert_allclose(shell, [0, 0, 3.0, 3.0], atol=1e-15)
    kx, ky, kz = D.spectra_1d(q, oper)
    assert kx[5] == 12.0 and kx.sum() == 12.0
    assert ky[0] == 6.0 and kz[0] == 6.0
    kzkh = D.spectrum_kzkh(q, oper)
    np.testing.assert_allclose(kzkh[0], [0, 0, 3.0, 3.0], atol=1e-15)
    # beyond the last shell: everything in the last bin
    q[:] = 0
    q[3, 0, 7] = 1.0  # |k| = sqrt(3.5^2 + 3^2) > 3
    np.testing.assert_allclose(D.shell_spectrum(q, oper), [0, 0, 0, 2.0], atol=1e-15)


def _order_pair(err_fn, dts=(2e-3, 1e-3)):
    e1, e2 = err_fn(dts[0]), err_fn(dts[1])
    return e1, e2, e1 / e2


def test_ns2d_budget_is_rate_of_change():
    oper = Oper2D(32, 32)
    rot = smooth_state(oper, np.random.default_rng(2), 1, 8)[0] * 20

    T = D.ns2d_budget(oper, rot)

    def err(dt):
        sim = NS2D(oper, type_time_scheme="Euler", deltat0=dt)
        sim.state_spect[0] = rot
        e0, z0 = D.energy_2d(oper, rot), np.abs(rot) ** 2 / 2
        sim.one_time_step_computation()
        r1 = sim.state_spect[0]
        dE = (D.shell_spectrum(D.energy_2d(oper, r1) - e0, oper)) / dt
        dZ = (D.shell_spectrum(np.abs(r1) ** 2 / 2 - z0, oper)) / dt
        scale = np.abs(T["transfer2D_E"]).max()
        return max(np.abs(dE - T["transfer2D_E"]).max() / scale,
                   np.abs(dZ - T["transfer2D_Z"]).max() / np.abs(T["transfer2D_Z"]).max())

    e1, e2, ratio = _order_pair(err)
    assert e1 < 0.2 and 1.7 < ratio < 2.3, (e1, e2, ratio)
    # alias-free: both transfers sum to zero
    assert abs(T["transfer2D_E"].sum()) < 1e-12 * np.abs(T["transfer2D_E"]).sum()
    assert abs(T["transfer2D_Z"].sum()) < 1e-12 * np.abs(T["transfer2D_Z"]).sum()


def test_ns3d_budget_is_rate_of_change():
    oper = Oper3D(16, 16, 16)
    S = smooth_state(oper, np.random.default_rng(3), 3, 5)
    oper.project_perpk3d(*S)
    f_d = np.zeros(oper.shapeK)
    T = D.ns3d_budget(oper, *S, f_d)
    _, _, kzkh0 = D.ns3d_spectra(oper, *S)
    s1d0, _, _ = D.ns3d_spectra(oper, *S)

    def err(dt):
        sim = NS3D(oper, type_time_scheme="Euler", deltat0=dt)
        sim.state_spect[:] = S
        sim.one_time_step_computation()
        s1d, _, kzkh = D.ns3d_spectra(oper, *sim.state_spect)
        dK = (kzkh["K"] - kzkh0["K"]) / dt
        dKhr = (kzkh["Khr"] - kzkh0["Khr"]) / dt
        dKz = (s1d["spectra_vz_kx"] - s1d0["spectra_vz_kx"]) / dt
        TK = T["transfer_Kh"] + T["transfer_Kz"]
        return max(np.abs(dK - TK).max() / np.abs(TK).max(),
                   np.abs(dKhr - T["transfer_Khr"]).max() / np.abs(T["transfer_Khr"]).max(),
                   np.abs(dKz - T["transfer_Kz_kx"]).max() / np.abs(T["transfer_Kz_kx"]).max())

    e1, e2, ratio = _order_pair(err)
    assert e1 < 0.2 and 1.7 < ratio < 2.3, (e1, e2, ratio)
    assert abs((T["transfer_Kh"] + T["transfer_Kz"]).sum()) < 1e-12 * np.abs(
        T["transfer_Kh"]).sum()


def test_strat_budget_is_rate_of_change():
    oper = Oper3D(16, 16, 16)
    N = 2.0
    S = smooth_state(oper, np.random.default_rng(4), 4, 5)
    oper.project_perpk3d(*S[:3])
    f_d = np.zeros(oper.shapeK)
    T = D.ns3d_budget(oper, *S[:3], f_d, b=S[3], N=N)

    def energies(S):
        eK = 0.5 * (np.abs(S[0]) ** 2 + np.abs(S[1]) ** 2 + np.abs(S[2]) ** 2)
        eA = 0.5 / N**2 * np.abs(S[3]) ** 2
        return D.spectrum_kzkh(eK, oper), D.spectrum_kzkh(eA, oper)

    K0, A0 = energies(S)
    TK = T["transfer_Kh"] + T["transfer_Kz"] - T["conv_K2A"]
    TA = T["transfer_A"] + T["conv_K2A"]

    def err(dt):
        sim = NS3DStrat(oper, N=N, type_time_scheme="Euler", deltat0=dt)
        sim.state_spect[:] = S
        sim.one_time_step_computation()
        K1, A1 = energies(sim.state_spect)
        return max(np.abs((K1 - K0) / dt - TK).max() / np.abs(TK).max(),
                   np.abs((A1 - A0) / dt - TA).max() / np.abs(TA).max())

    e1, e2, ratio = _order_pair(err)
    assert e1 < 0.2 and 1.7 < ratio < 2.3, (e1, e2, ratio)
    assert abs(T["transfer_A"].sum()) < 1e-12 * np.abs(T["transfer_A"]).sum()


def test_dissipation_terms_are_viscous_decay():
    oper = Oper3D(16, 16, 16)
    N = 1.5
    S = smooth_state(oper, np.random.default_rng(5), 4, 6)
    sim = NS3DStrat(oper, N=N, nu_2=0.02, nu_4=1e-4)
    f_d, _ = sim.compute_freq_diss()
    T = D.ns3d_budget(oper, *S[:3], f_d, b=S[3], N=N)
    for name, idx, coef in (("diss_Kz", [2], 0.5), ("diss_A", [3], 0.5 / N**2)):
        errs = []
        for dt in (2e-3, 1e-3):
            S1 = S * np.exp(-f_d * dt)
            e0 = sum(coef * np.abs(S[i]) ** 2 for i in idx)
            e1 = sum(coef * np.abs(S1[i]) ** 2 for i in idx)
            rate = D.spectrum_kzkh((e1 - e0) / dt, oper)
            errs.append(np.abs(rate + T[name]).max() / np.abs(T[name]).max())
        assert errs[0] < 1e-2 and 1.7 < errs[0] / errs[1] < 2.3, (name, errs)


def test_means_injection_and_energy():
    np.random.seed(20)
    o2 = Oper2D(32, 32)
    rot = smooth_state(o2, np.random.default_rng(6), 1, 8)[0]
    forc = TCRandomForcing(o2, ("rot_fft",), 3, 5, forcing_rate=0.3)
    dt = 0.01
    Frot = forc.compute(np.array([rot]), 0.0, dt)[0]
    sim = NS2D(o2, nu_2=1e-3)
    f_d, f_d_hypo = sim.compute_freq_diss()
    m = D.ns2d_means(o2, rot, f_d, f_d_hypo, Frot, dt)
    assert m["PZ1"] + m["PZ2"] == pytest.approx(0.3, rel=1e-12)
    d1, d2 = D.ns2d_spectra(o2, rot)
    assert d2["spectrum2D_E"].sum() * o2.deltakx == pytest.approx(m["E"], rel=1e-13)

    np.random.seed(21)
    o3 = Oper3D(16, 16, 16)
    S = smooth_state(o3, np.random.default_rng(7), 3, 5)
    forc = TCRandomForcing(o3, ("vx_fft", "vy_fft", "vz_fft"), 2, 3, key_forced="vx_fft")
    f = forc.compute(S, 0.0, dt)
    sim = NS3D(o3, nu_2=1e-3)
    f_d, f_d_hypo = sim.compute_freq_diss()
    m = D.ns3d_means(o3, *S, f_d, f_d_hypo, f=f, deltat=dt)
    assert m["PK1"] + m["PK2"] == pytest.approx(1.0, rel=1e-12)
    _, d3, _ = D.ns3d_spectra(o3, *S)
    assert d3["spectra_E"].sum() * D.Bins(o3).dk == pytest.approx(m["E"], rel=1e-13)


def test_planted_mistake_weights_dropped(monkeypatch):
    oper = Oper3D(8, 6, 4)
    X = np.random.default_rng(8).standard_normal(oper.shapeX)
    q = np.abs(oper.fft(X)) ** 2 / 2
    ref = brute_force(oper, X)
    monkeypatch.setattr(D, "parseval_weights", lambda o: np.ones(o.shapeK))
    assert not np.allclose(D.spectra_1d(q, oper)[0], ref["kx"], rtol=1e-3)


def test_planted_mistake_transfer_sign():
    oper = Oper2D(32, 32)
    rot = smooth_state(oper, np.random.default_rng(9), 1, 8)[0] * 20
    T = D.ns2d_budget(oper, rot)
    dt = 1e-3
    sim = NS2D(oper, type_time_scheme="Euler", deltat0=dt)
    sim.state_spect[0] = rot
    e0 = D.energy_2d(oper, rot)
    sim.one_time_step_computation()
    dE = D.shell_spectrum(D.energy_2d(oper, sim.state_spect[0]) - e0, oper) / dt
    bad = -T["transfer2D_E"]
    assert np.abs(dE - bad).max() > 0.5 * np.abs(bad).max()
