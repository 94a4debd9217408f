"""Pins of the resolution-change oracle (oracle/resize.py).

* a field band-limited inside the coarse grid keeps its physical values:
  the coarse inverse transform of the truncated spectrum equals the fine
  field sampled on the coarse points (and refinement back is exact);
* refine-then-truncate is the identity on the coarse spectrum;
* a mistake (negative wavenumbers not copied) breaks the sampling pin.
"""

import numpy as np

from oracle.operators import Oper2D, Oper3D
from oracle import resize as R


def band_limited(o, seed, kmax):
    rng = np.random.default_rng(seed)
    a = o.fft(rng.standard_normal(o.shapeX))
    a[np.sqrt(o.K2) > kmax] = 0
    return o.fft(o.ifft(a))


def test_2d_down_keeps_values_and_up_is_exact():
    fine, coarse = Oper2D(64, 32), Oper2D(16, 16)
    a = band_limited(fine, 0, 6)
    c = np.zeros(coarse.shapeK, complex)
    R.fill_field_fft_2d(a, c)
    x_fine = fine.ifft(a)
    np.testing.assert_allclose(coarse.ifft(c), x_fine[::2, ::4], atol=1e-12)
    back = np.zeros(fine.shapeK, complex)
    R.fill_field_fft_2d(c, back)
    np.testing.assert_allclose(back, a, atol=1e-14)


def test_3d_down_keeps_values_and_roundtrip():
    fine, coarse = Oper3D(32, 32, 16), Oper3D(16, 8, 8)
    a = band_limited(fine, 1, 3.5)
    c = np.zeros(coarse.shapeK, complex)
    R.fill_field_fft_3d(a, c)
    np.testing.assert_allclose(coarse.ifft(c), fine.ifft(a)[::2, ::4, ::2], atol=1e-12)
    back = np.zeros(fine.shapeK, complex)
    R.fill_field_fft_3d(c, back)
    c2 = np.zeros(coarse.shapeK, complex)
    R.fill_field_fft_3d(back, c2)
    np.testing.assert_allclose(c2, c, atol=1e-15)


def test_planted_mistake():
    fine, coarse = Oper2D(32, 32), Oper2D(16, 16)
    a = band_limited(fine, 2, 5)
    c = np.zeros(coarse.shapeK, complex)
    nk0 = 16
    for ik0 in range(nk0 // 2 + 1):  # positive ky only
        c[ik0, :9] = a[ik0, :9]
    assert np.abs(coarse.ifft(c) - fine.ifft(a)[::2, ::2]).max() > 1e-3
