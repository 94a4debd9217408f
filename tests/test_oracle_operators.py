"""Pins of the oracle's operators (CPU only).

Each check is chosen so that a plausible mistake (wrong normalisation, sign,
index map, transposed operand) fails it; none retypes the oracle's formula.
"""

import itertools

import numpy as np
import pytest

from oracle.operators import Oper2D, Oper3D


def brute_dft(x):
    """Forward DFT by direct summation, divided by the number of points."""
    x = np.asarray(x, dtype=np.float64)
    shape = x.shape
    out = np.zeros(shape, dtype=np.complex128)
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    for kidx in itertools.product(*[range(n) for n in shape]):
        phase = sum(k * g / n for k, g, n in zip(kidx, grids, shape))
        out[kidx] = np.sum(x * np.exp(-2j * np.pi * phase))
    out /= x.size
    return out[..., : shape[-1] // 2 + 1]


@pytest.mark.parametrize("shape", [(6, 8), (5, 7), (4, 6, 8), (3, 5, 4)])
def test_fft_matches_brute_force_dft(shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape)
    if len(shape) == 2:
        oper = Oper2D(shape[1], shape[0])
    else:
        oper = Oper3D(shape[2], shape[1], shape[0])
    np.testing.assert_allclose(oper.fft(x), brute_dft(x), rtol=0, atol=1e-14)
    np.testing.assert_allclose(oper.ifft(oper.fft(x)), x, rtol=0, atol=1e-13)


@pytest.mark.parametrize("nx", [16, 15])
def test_parseval_weights(nx):
    """sum_wavenumbers(|f|^2/2) == mean(f^2/2) (fluidsim/base/state.py:385-392)
    pins both the 1/N normalisation and the half-spectrum weights."""
    rng = np.random.default_rng(1)
    oper = Oper2D(nx, 12)
    x = rng.standard_normal(oper.shapeX)
    E_spect = oper.sum_wavenumbers(np.abs(oper.fft(x)) ** 2 / 2)
    assert E_spect == pytest.approx(np.mean(x**2) / 2, rel=1e-14)
    oper3 = Oper3D(nx, 6, 4)
    x3 = rng.standard_normal(oper3.shapeX)
    E3 = oper3.sum_wavenumbers(np.abs(oper3.fft(x3)) ** 2 / 2)
    assert E3 == pytest.approx(np.mean(x3**2) / 2, rel=1e-14)


def test_single_mode_closed_forms_2d():
    """cos(3x + 2y) on a 2pi box: modes (ky, kx) = (2, 3) and (-2, -3) hold
    1/2; the curl/velocity/gradient of a single mode are known in closed
    form."""
    oper = Oper2D(16, 16)
    y, x = np.meshgrid(
        np.arange(16) * oper.deltay, np.arange(16) * oper.deltax, indexing="ij"
    )
    f = np.cos(3 * x + 2 * y)
    f_fft = oper.fft(f)
    assert f_fft[2, 3] == pytest.approx(0.5, abs=1e-15)
    assert f_fft[-2, 3] == pytest.approx(0.0, abs=1e-15)
    px, py = oper.gradfft_from_fft(f_fft)
    np.testing.assert_allclose(oper.ifft(px), -3 * np.sin(3 * x + 2 * y), atol=1e-13)
    np.testing.assert_allclose(oper.ifft(py), -2 * np.sin(3 * x + 2 * y), atol=1e-13)
    # rot = cos(3x+2y) -> psi = -rot/k^2 with u = -dpsi/dy, v = dpsi/dx
    # (u = d_y(rot)/k^2 ... ) closed form: u = -2 sin(.)/13, v = 3 sin(.)/13
    ux_fft, uy_fft = oper.vecfft_from_rotfft(f_fft)
    s = np.sin(3 * x + 2 * y)
    np.testing.assert_allclose(oper.ifft(ux_fft), -2 * s / 13, atol=1e-14)
    np.testing.assert_allclose(oper.ifft(uy_fft), 3 * s / 13, atol=1e-14)


def test_curl_round_trip_2d():
    """fluidsim/operators/test/test_operators2d.py:69-80 (rtol 1e-15,
    atol 1e-14) on the odd grid the reference uses (nh=9) and on 16."""
    rng = np.random.default_rng(2)
    for nh in (9, 16):
        oper = Oper2D(nh, nh, 6.0, 6.0)
        rot_fft = oper.fft(rng.standard_normal(oper.shapeX))
        rot_fft[0, 0] = 0.0
        ux_fft, uy_fft = oper.vecfft_from_rotfft(rot_fft)
        rot2_fft = oper.rotfft_from_vecfft(ux_fft, uy_fft)
        np.testing.assert_allclose(rot2_fft, rot_fft, 1e-15, 1e-14)
        # velocity is divergence free
        div = oper.divfft_from_vecfft(ux_fft, uy_fft)
        assert np.abs(div).max() < 1e-14


def test_laplacian_2d():
    """test_operators2d.py:82-100: K2 * f exact, inverse through K2_not0."""
    rng = np.random.default_rng(3)
    oper = Oper2D(9, 9, 6.0, 6.0)
    ff_fft = oper.fft(rng.standard_normal(oper.shapeX))
    ff_fft[0, 0] = 0.0
    lap = oper.K2 * ff_fft
    back = lap / oper.K2_not0
    np.testing.assert_allclose(back, ff_fft, 1e-15, 1e-14)
    # Laplacian of cos(2x) is -4 cos(2x) on a 2pi box
    op2 = Oper2D(8, 8)
    x = np.arange(8) * op2.deltax
    f = np.cos(2 * x)[None, :] * np.ones((8, 1))
    np.testing.assert_allclose(
        op2.ifft(-op2.K2 * op2.fft(f)), -4 * f, atol=1e-13
    )


def test_projection_perp_2d():
    rng = np.random.default_rng(4)
    oper = Oper2D(12, 10)
    ux = oper.fft(rng.standard_normal(oper.shapeX))
    uy = oper.fft(rng.standard_normal(oper.shapeX))
    oper.projection_perp(ux, uy)
    assert np.abs(oper.KX * ux + oper.KY * uy).max() < 1e-14
    ux2, uy2 = ux.copy(), uy.copy()
    oper.projection_perp(ux2, uy2)
    np.testing.assert_allclose(ux2, ux, atol=1e-15)


def _rand_vec_fft(oper, rng):
    return [oper.fft(rng.standard_normal(oper.shapeX)) for _ in range(3)]


def test_projections_3d():
    """radial + poloidal + toroidal = identity and cross projections vanish
    (fluidsim/operators/test/test_operators3d.py:353-480), plus k.P(v) = 0
    and idempotence of project_perpk3d."""
    rng = np.random.default_rng(5)
    oper = Oper3D(8, 6, 4, 2 * np.pi, 3.0, 5.0)
    v = _rand_vec_fft(oper, rng)
    for a in v:
        a.flat[0] = 0.0
    # the decomposition uses angles; it is exact where kh != 0
    kh0 = (oper.Kx**2 + oper.Ky**2) == 0
    parts = []
    for proj in (oper.project_kradial3d, oper.project_poloidal, oper.project_toroidal):
        w = [a.copy() for a in v]
        proj(*w)
        parts.append(w)
    for i in range(3):
        total = parts[0][i] + parts[1][i] + parts[2][i]
        np.testing.assert_allclose(total[~kh0], v[i][~kh0], atol=1e-14)
    # projections are orthogonal: toroidal part has no radial component
    w = [a.copy() for a in parts[2]]
    oper.project_kradial3d(*w)
    assert max(np.abs(a).max() for a in w) < 1e-14
    # Leray projection
    p = [a.copy() for a in v]
    oper.project_perpk3d(*p)
    kdotv = oper.Kx * p[0] + oper.Ky * p[1] + oper.Kz * p[2]
    assert np.abs(kdotv).max() < 1e-13
    p2 = [a.copy() for a in p]
    oper.project_perpk3d(*p2)
    for a, b in zip(p, p2):
        np.testing.assert_allclose(a, b, atol=1e-15)
    # toroidal component round trip (operators3d.py:939-972)
    vt = oper.vtfft_from_vecfft(*parts[2])
    back = oper.vecfft_from_vtfft(vt)
    for i in range(2):
        np.testing.assert_allclose(back[i][~kh0], parts[2][i][~kh0], atol=1e-14)


def test_curl_3d_single_mode():
    """v = (0, 0, cos(2x)) -> omega = curl v = (0, 2 sin(2x), 0)."""
    oper = Oper3D(8, 8, 8)
    x = np.arange(8) * oper.deltax
    vz = np.broadcast_to(np.cos(2 * x), oper.shapeX).copy()
    zero = np.zeros(oper.shapeK, dtype=complex)
    om = oper.rotfft_from_vecfft(zero, zero, oper.fft(vz))
    np.testing.assert_allclose(oper.ifft(om[0]), 0, atol=1e-14)
    np.testing.assert_allclose(
        oper.ifft(om[1]), np.broadcast_to(2 * np.sin(2 * x), oper.shapeX), atol=1e-13
    )
    np.testing.assert_allclose(oper.ifft(om[2]), 0, atol=1e-14)


@pytest.mark.parametrize("n, c", [(12, 2 / 3), (16, 2 / 3), (256, 2 / 3), (10, 0.5)])
def test_dealias_mask_cubic_brute_force(n, c):
    """A mode is dealiased iff |m| >= c n / 2 along some axis, m the signed
    integer wavenumber with the Nyquist index at +n/2."""
    oper = Oper2D(n, n, 6.0, 6.0, coef_dealiasing=c)
    mask = oper.where_dealiased
    for iy in range(n):
        my = iy if iy <= n // 2 else iy - n
        for ix in range(n // 2 + 1):
            expect = (abs(ix) >= c * n / 2) or (abs(my) >= c * n / 2)
            assert bool(mask[iy, ix]) == expect


def test_dealias_false_is_identity():
    """test_operators2d.py:138-154."""
    oper = Oper2D(9, 9, 6.0, 6.0, coef_dealiasing=False)
    a = oper.create_arrayK(1.0)
    s = a.sum()
    oper.dealiasing(a)
    assert a.sum() == s


def test_dealias_spherical_and_no_multiple_aliases():
    oper = Oper2D(16, 16, truncation_shape="spherical")
    kmax = 2 / 3 * 8
    np.testing.assert_array_equal(oper.where_dealiased, oper.K2 >= kmax**2)
    oper = Oper2D(16, 16, truncation_shape="no_multiple_aliases")
    al = (np.abs(oper.KX) >= 2 / 3 * 8) & (np.abs(oper.KY) >= 2 / 3 * 8)
    np.testing.assert_array_equal(
        oper.where_dealiased.astype(bool), al | (oper.K2 >= kmax**2)
    )
