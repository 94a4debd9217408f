"""FP64 restatement of the convolutional closure network.

Test infrastructure only (see ``oracle/__init__.py``).

No reference counterpart: fluidsim has no learned closure (its LES hook is
``TurbModel.get_forcing``, fluidsim/base/turb_model/base.py:87-88, with
Smagorinsky only).  The network follows the description of the closure
workload: 5x5 convolutions with periodic padding, 32 channels, GELU
(exact, erf form) between layers, mean-squared-error loss.

Layouts match the engine (channels last): x[B][H][W][C], weights
w[O][25][C] with tap t = 5 dy + dx and offset (dy - 2, dx - 2), bias b[O].

  forward      y[p, o] = b[o] + sum_{t, c} w[o, t, c] x[p + d(t), c]
  input grad   dx[p, c] = sum_{t, o} w[o, t, c] dy[p - d(t), o]
  weight grad  dw[o, t, c] = sum_p dy[p, o] x[p + d(t), c],  db = sum_p dy

Pinned by (tests/test_oracle_cnn.py): the convolution equals the FFT form
of a periodic cross-correlation; directional finite differences of the
loss match the analytic gradients of every weight, bias and the input;
an unflipped kernel in the input gradient fails the check.
"""

import numpy as np
from scipy.special import erf

TAPS = [(t // 5 - 2, t % 5 - 2) for t in range(25)]


def shift(x, d):
    """x[p + d] with periodic wrap over the (H, W) axes of (B, H, W, C)."""
    return np.roll(x, (-d[0], -d[1]), axis=(1, 2))


def conv(x, w, b, co=None):
    co = w.shape[0] if co is None else co
    y = np.zeros(x.shape[:3] + (co,))
    for t, d in enumerate(TAPS):
        y += shift(x, d) @ w[:co, t, : x.shape[3]].T
    return y + b[:co]


def conv_input_grad(dy, w, ci):
    co = dy.shape[3]
    dx = np.zeros(dy.shape[:3] + (ci,))
    for t, d in enumerate(TAPS):
        dx += shift(dy, (-d[0], -d[1])) @ w[:co, t, :ci]
    return dx


def conv_weight_grad(x, dy):
    co, ci = dy.shape[3], x.shape[3]
    dw = np.zeros((co, 25, ci))
    for t, d in enumerate(TAPS):
        xs = shift(x, d).reshape(-1, ci)
        dw[:, t, :] = dy.reshape(-1, co).T @ xs
    return dw, dy.reshape(-1, co).sum(axis=0)


def gelu(v):
    return 0.5 * v * (1 + erf(v / np.sqrt(2)))


def gelu_grad(v):
    return 0.5 * (1 + erf(v / np.sqrt(2))) + v * np.exp(-0.5 * v * v) / np.sqrt(2 * np.pi)


class Net:
    """layers: list of (w (O, 25, C), b (O,)); GELU after all but the last."""

    def __init__(self, layers):
        self.layers = layers

    def forward(self, x, keep=False):
        acts, pres = [x], []
        a = x
        for i, (w, b) in enumerate(self.layers):
            y = conv(a, w, b)
            pres.append(y)
            a = gelu(y) if i < len(self.layers) - 1 else y
            acts.append(a)
        return (a, acts, pres) if keep else a

    def loss(self, x, target):
        out = self.forward(x)
        return np.mean((out[..., 0] - target) ** 2)

    def gradients(self, x, target):
        """(loss, [(dw, db)], dx) of the mean-squared error."""
        out, acts, pres = self.forward(x, keep=True)
        n = target.size
        g = np.zeros_like(out)
        g[..., 0] = 2 * (out[..., 0] - target) / n
        grads = [None] * len(self.layers)
        for i in range(len(self.layers) - 1, -1, -1):
            w, b = self.layers[i]
            if i < len(self.layers) - 1:
                g = g * gelu_grad(pres[i])
            grads[i] = conv_weight_grad(acts[i], g)
            g = conv_input_grad(g, w, acts[i].shape[3])
        loss = np.mean((out[..., 0] - target) ** 2)
        return loss, grads, g
