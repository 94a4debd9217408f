"""Pins of the closure-network oracle (oracle/cnn.py).

* the periodic 5x5 convolution equals the FFT form of a periodic
  cross-correlation (an independent formulation);
* directional finite differences of the loss (central, FP64) match the
  analytic gradients of every weight, bias and the input to 1e-7 relative;
* a planted mistake (kernel not flipped in the input gradient) fails.
"""

import numpy as np
import pytest

from oracle import cnn as C


def small_net(rng, cin=3, hidden=4, nl=3, cout=1):
    layers = []
    c = cin
    for i in range(nl):
        o = cout if i == nl - 1 else hidden
        layers.append((rng.standard_normal((o, 25, c)) * 0.3, rng.standard_normal(o) * 0.1))
        c = o
    return C.Net(layers)


def test_conv_equals_fft_correlation():
    rng = np.random.default_rng(0)
    B, H, W, ci, co = 2, 8, 16, 3, 4
    x = rng.standard_normal((B, H, W, ci))
    w = rng.standard_normal((co, 25, ci))
    b = rng.standard_normal(co)
    y = C.conv(x, w, b)
    # kernel on the grid: k[o, c][(-dy, -dx) mod (H, W)] = w so that
    # y = ifft(fft(x) * fft(k)) is a correlation with offsets +d
    k = np.zeros((co, ci, H, W))
    for t, (dy, dx) in enumerate(C.TAPS):
        k[:, :, (-dy) % H, (-dx) % W] = w[:, t, :]
    X = np.fft.fft2(np.moveaxis(x, 3, 1))            # (B, ci, H, W)
    K = np.fft.fft2(k)                               # (co, ci, H, W)
    Y = np.einsum("bchw,ochw->bohw", X, K)
    yf = np.moveaxis(np.fft.ifft2(Y).real, 1, 3) + b
    np.testing.assert_allclose(y, yf, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    net = small_net(rng)
    x = rng.standard_normal((2, 8, 8, 3))
    target = rng.standard_normal((2, 8, 8))
    loss, grads, dx = net.gradients(x, target)
    assert loss == pytest.approx(net.loss(x, target), rel=1e-15)
    eps = 1e-6
    for i, (w, b) in enumerate(net.layers):
        for arr, g in ((w, grads[i][0]), (b, grads[i][1])):
            d = rng.standard_normal(arr.shape)
            arr += eps * d
            lp = net.loss(x, target)
            arr -= 2 * eps * d
            lm = net.loss(x, target)
            arr += eps * d
            fd = (lp - lm) / (2 * eps)
            an = np.sum(g * d)
            assert an == pytest.approx(fd, rel=1e-7, abs=1e-12), (i, arr.shape)
    d = rng.standard_normal(x.shape)
    fd = (net.loss(x + eps * d, target) - net.loss(x - eps * d, target)) / (2 * eps)
    assert np.sum(dx * d) == pytest.approx(fd, rel=1e-7)


def test_planted_mistake_unflipped_kernel(monkeypatch):
    rng = np.random.default_rng(2)
    net = small_net(rng)
    x = rng.standard_normal((1, 8, 8, 3))
    target = rng.standard_normal((1, 8, 8))

    def bad_input_grad(dy, w, ci):
        dx = np.zeros(dy.shape[:3] + (ci,))
        for t, d in enumerate(C.TAPS):
            dx += C.shift(dy, d) @ w[: dy.shape[3], t, :ci]  # shift not reversed
        return dx

    monkeypatch.setattr(C, "conv_input_grad", bad_input_grad)
    _, _, dx = net.gradients(x, target)
    eps = 1e-6
    d = rng.standard_normal(x.shape)
    fd = (net.loss(x + eps * d, target) - net.loss(x - eps * d, target)) / (2 * eps)
    assert abs(np.sum(dx * d) - fd) > 1e-3 * abs(fd)


def test_training_symmetries_of_the_ns2d_tendency():
    """The closure training turns and mirrors its samples: the ns2d nonlinear
    tendency must commute with a 90-degree turn of a square periodic box and
    with a mirror that flips the sign of the vorticity (pseudoscalar)."""
    from oracle.operators import Oper2D
    from oracle.solvers import NS2D

    n = 32
    oper = Oper2D(n, n)
    sol = NS2D(oper)
    rng = np.random.default_rng(3)
    w_fft = oper.fft(rng.standard_normal(oper.shapeX))
    oper.dealiasing(w_fft)
    w = oper.ifft(w_fft)

    def tend(wp):
        return oper.ifft(sol.tendencies_nonlin(np.array([oper.fft(wp)]))[0])

    N = tend(w)
    scale = np.abs(N).max()
    for turn in (1, 2, 3):
        got = tend(np.rot90(w, turn).copy())
        assert np.abs(got - np.rot90(N, turn)).max() < 1e-12 * scale
    got = tend(-np.flip(w, 1).copy())
    assert np.abs(got + np.flip(N, 1)).max() < 1e-12 * scale
    # a mirror without the sign change is not a symmetry
    bad = tend(np.flip(w, 1).copy())
    assert np.abs(bad - np.flip(N, 1)).max() > 1e-3 * scale
