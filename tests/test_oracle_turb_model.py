"""Pins of the Smagorinsky oracle (oracle/turb_model.py).

* exact energy identity: with g_ij = 2 C_nuT |S| S_ij and f = 2 div g, the
  Parseval sum of Re(conj(v) f) equals -4 C_nuT mean(|S|^3) (discrete
  integration by parts of spectral derivatives of a band-limited field);
* the strain rate of a solenoidal field is traceless and S_ij of
  v = (sin z, 0, 0) is the closed form S_xz = cos(z)/2;
* brute force: f_x of a tiny field from finite sums of the definitions;
* planted mistakes (dropped 2, wrong sign) break the energy identity.
"""

import numpy as np
import pytest

from oracle.operators import Oper3D
from oracle.turb_model import Smagorinsky


def solenoidal(o, seed, kmax=4):
    rng = np.random.default_rng(seed)
    v = [o.fft(rng.standard_normal(o.shapeX)) for _ in range(3)]
    for a in v:
        a[o.K > kmax] = 0
    o.project_perpk3d(*v)
    return v


def energy_rate(o, v, f):
    return sum(o.sum_wavenumbers((a.conj() * b).real) for a, b in zip(v, f))


def test_energy_identity():
    o = Oper3D(24, 24, 24)
    m = Smagorinsky(o, C=0.18)
    v = solenoidal(o, 0)
    f = m.get_forcing(*v)
    S = m.stress_tensor(*v)
    normS = m.norm(*S)
    expected = -4 * m.C_nu_T * np.mean(normS**3)
    assert energy_rate(o, v, f) == pytest.approx(expected, rel=1e-12)
    assert expected < 0


def test_traceless_and_shear_closed_form():
    o = Oper3D(16, 16, 16)
    m = Smagorinsky(o)
    v = solenoidal(o, 1)
    Sxx, Syy, Szz, *_ = m.stress_tensor(*v)
    assert np.abs(Sxx + Syy + Szz).max() < 1e-12 * np.abs(Sxx).max()
    z = np.arange(16) * 2 * np.pi / 16
    vx = np.broadcast_to(np.sin(z)[:, None, None], o.shapeX)
    zero = np.zeros(o.shapeX)
    S = m.stress_tensor(o.fft(vx), o.fft(zero), o.fft(zero))
    np.testing.assert_allclose(S[4], 0.5 * np.broadcast_to(np.cos(z)[:, None, None], o.shapeX), atol=1e-14)
    for i in (0, 1, 2, 3, 5):
        assert np.abs(S[i]).max() < 1e-14


def test_brute_force_fx_tiny():
    n = 4
    o = Oper3D(n, n, n)
    m = Smagorinsky(o, C=0.3)
    v = solenoidal(o, 2, kmax=1.5)
    fx = m.get_forcing(*v)[0]
    # brute-force DFTs of the definitions
    grid = np.arange(n) * 2 * np.pi / n
    Z, Y, X = np.meshgrid(grid, grid, grid, indexing="ij")
    k1 = np.fft.fftfreq(n, 1 / n)
    phys = [np.zeros(o.shapeX) for _ in range(3)]
    grads = [[np.zeros(o.shapeX) for _ in range(3)] for _ in range(3)]
    full = [np.fft.fftn(o.ifft(a), norm="forward") for a in v]
    for iz in range(n):
        for iy in range(n):
            for ix in range(n):
                kz, ky, kx = k1[iz], k1[iy], k1[ix]
                e = np.exp(1j * (kx * X + ky * Y + kz * Z))
                for c in range(3):
                    a = full[c][iz, iy, ix]
                    phys[c] += (a * e).real
                    for d, kd in enumerate((kx, ky, kz)):
                        grads[c][d] += (1j * kd * a * e).real
    (dx_ux, dy_ux, dz_ux), (dx_uy, dy_uy, dz_uy), (dx_uz, dy_uz, dz_uz) = grads
    Sxx, Syx, Szx = dx_ux, 0.5 * (dy_ux + dx_uy), 0.5 * (dz_ux + dx_uz)
    Syy, Szz, Szy = dy_uy, dz_uz, 0.5 * (dz_uy + dy_uz)
    nrm = np.sqrt(Sxx**2 + Syy**2 + Szz**2 + 2 * (Syx**2 + Szx**2 + Szy**2))
    g = [2 * m.C_nu_T * nrm * s for s in (Sxx, Syx, Szx)]
    ref = np.zeros(o.shapeK, complex)
    for iz in range(n):
        for iy in range(n):
            for ix in range(n // 2 + 1):
                kz, ky, kx = k1[iz], k1[iy], k1[ix] if ix < n // 2 else n // 2
                e = np.exp(-1j * (kx * X + ky * Y + kz * Z))
                gh = [np.mean(gg * e) for gg in g]
                ref[iz, iy, ix] = 2j * (o.Kx[iz, iy, ix] * gh[0] + o.Ky[iz, iy, ix] * gh[1]
                                        + o.Kz[iz, iy, ix] * gh[2])
    np.testing.assert_allclose(fx, ref, atol=1e-13)


def test_planted_mistakes_break_identity():
    o = Oper3D(16, 16, 16)
    m = Smagorinsky(o)
    v = solenoidal(o, 3)
    S = m.stress_tensor(*v)
    expected = -4 * m.C_nu_T * np.mean(m.norm(*S) ** 3)
    f = m.get_forcing(*v)
    half = [a / 2 for a in f]  # dropped factor 2
    flip = [-a for a in f]     # wrong sign
    for bad in (half, flip):
        assert abs(energy_rate(o, v, bad) - expected) > 0.1 * abs(expected)
