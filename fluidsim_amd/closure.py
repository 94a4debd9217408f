"""Learned subgrid closure for coarse 2D LES (ns2d), on the matrix cores.

fluidsim has no learned closure: its LES hook is ``TurbModel.get_forcing``
(fluidsim/base/turb_model/base.py:87-88, Smagorinsky only, ns3d only); ns2d
calls no model (fluidsim/solvers/ns2d/solver.py:191-192).  This module adds

* ``CNNClosure``: 5x5 periodic convolutions (32 channels, GELU), BF16
  weights and activations with FP32 accumulation on MFMA (``fsx_cnn_*``),
  FP32 master weights trained with Adam on the device;
* ``subgrid_terms``: training pairs from a DNS state, filtered by the sharp
  spectral cut of fluidsim's resolution change (fill_field_fft_2d,
  fluidsim/base/init_fields.py:300-314): input the coarse vorticity, target
  the subgrid vorticity tendency S = cut(N(w)) - N(cut(w)) with N the ns2d
  nonlinear term (fluidsim/solvers/ns2d/solver.py:111-194);
* ``SimulLES``: the ns2d solver on the coarse grid with the network's S
  added to the nonlinear tendency at every Runge-Kutta stage.
"""

import ctypes

import numpy as np

from . import _lib


def _torch():
    import torch

    return torch


def bf16_bits(a):
    """Round-to-nearest-even BF16 bit patterns (int16 array) of float data."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r.view(np.int16)


def bf16_value(bits):
    return (np.asarray(bits).view(np.uint16).astype(np.uint32) << 16).view(np.float32)


class CNNClosure:
    """Convolutional network x (B, H, W, n_in) -> y (B, H, W) on the device.

    Layer l maps cin[l] (8 for the input layer, channel-padded, else 32)
    to cout[l] channels (32, the last layer 1) with a periodic 5x5
    convolution; GELU after every layer but the last.
    """

    def __init__(self, n_in=1, hidden=32, nlayers=6, seed=0, lr=1e-3):
        torch = _torch()
        if hidden != 32:
            raise ValueError("the MFMA kernels use 32 hidden channels")
        if not 1 <= n_in <= 8:
            raise ValueError("1 to 8 input channels")
        self.n_in = n_in
        self.nlayers = nlayers
        self.cin = [8] + [32] * (nlayers - 1)
        self.cout = [32] * (nlayers - 1) + [1]
        self.lr = lr
        rng = np.random.default_rng(seed)
        self.w, self.b = [], []
        for l in range(nlayers):
            fan_in = 25 * (n_in if l == 0 else 32)
            w = np.zeros((32, 25, self.cin[l]), np.float32)
            std = np.sqrt(2.0 / fan_in) * (0.1 if l == nlayers - 1 else 1.0)
            real_in = n_in if l == 0 else 32
            w[: self.cout[l], :, :real_in] = rng.standard_normal(
                (self.cout[l], 25, real_in)) * std
            self.w.append(torch.from_numpy(w).cuda())
            self.b.append(torch.zeros(32, dtype=torch.float32, device="cuda"))
        self.wb = [torch.empty(w.shape, dtype=torch.int16, device="cuda") for w in self.w]
        self._refresh_bf16()
        self.m1 = [torch.zeros_like(t) for t in self.w + self.b]
        self.m2 = [torch.zeros_like(t) for t in self.w + self.b]
        self.step_count = 0
        self._work = {}
        self.zero_bias = torch.zeros(32, dtype=torch.float32, device="cuda")

    # ------------------------------------------------------------ state
    def _refresh_bf16(self):
        for w, wb in zip(self.w, self.wb):
            _lib.call("fsx_cnn_to_bf16", _lib.dptr(w), w.numel(), _lib.dptr(wb),
                      _lib.stream_ptr())

    def state_dict(self):
        return {f"w{l}": self.w[l].cpu().numpy() for l in range(self.nlayers)} | {
            f"b{l}": self.b[l].cpu().numpy() for l in range(self.nlayers)}

    def load_state_dict(self, d):
        torch = _torch()
        for l in range(self.nlayers):
            self.w[l].copy_(torch.from_numpy(np.asarray(d[f"w{l}"], np.float32)))
            self.b[l].copy_(torch.from_numpy(np.asarray(d[f"b{l}"], np.float32)))
        self._refresh_bf16()

    # ------------------------------------------------------------ buffers
    def _buffers(self, B, H, W):
        torch = _torch()
        key = (B, H, W)
        if key not in self._work:
            npix = B * H * W
            self._work[key] = {
                "x": torch.empty((B, H, W, 8), dtype=torch.int16, device="cuda"),
                "fwd": torch.empty(2 * npix * 32 * 2 + npix * 32 * 4, dtype=torch.uint8,
                                   device="cuda"),
            }
        return self._work[key]

    def pack_input(self, x):
        """(B, H, W, n_in) float -> BF16 (B, H, W, 8) channel-padded."""
        B, H, W, c = x.shape
        buf = self._buffers(B, H, W)["x"]
        xc = x.contiguous()  # held until the launch is issued
        _lib.call("fsx_cnn_act", _lib.dptr(xc), B * H * W, c, 8, 0, _lib.dptr(buf),
                  _lib.stream_ptr())
        return buf

    # ------------------------------------------------------------ inference
    def forward(self, x, out=None):
        """Network output (B, H, W) float32 for x (B, H, W, n_in) float32."""
        torch = _torch()
        B, H, W, _ = x.shape
        xb = self.pack_input(x)
        work = self._buffers(B, H, W)["fwd"]
        if out is None:
            out = torch.empty((B, H, W), dtype=torch.float32, device="cuda")
        nl = self.nlayers
        cin = (ctypes.c_int * nl)(*self.cin)
        cout = (ctypes.c_int * nl)(*self.cout)
        wp = (ctypes.c_void_p * nl)(*[t.data_ptr() for t in self.wb])
        bp = (ctypes.c_void_p * nl)(*[t.data_ptr() for t in self.b])
        _lib.call("fsx_cnn_forward", nl, cin, cout, wp, bp, _lib.dptr(xb), B, H, W,
                  _lib.dptr(out), _lib.dptr(work), work.numel(), _lib.stream_ptr())
        return out

    # ------------------------------------------------------------ training
    def train_step(self, x, target, lr=None, want_loss=True):
        """One Adam step on the mean-squared error of x -> target."""
        torch = _torch()
        s = _lib.stream_ptr()
        B, H, W, _ = x.shape
        npix = B * H * W
        dev = dict(dtype=torch.float32, device="cuda")
        acts = [self.pack_input(x)]
        pres = []
        for l in range(self.nlayers):
            last = l == self.nlayers - 1
            y = torch.empty((B, H, W, self.cout[l]), **dev)
            _lib.call("fsx_cnn_conv_fwd", _lib.dptr(acts[l]), B, H, W, self.cin[l],
                      _lib.dptr(self.wb[l]), _lib.dptr(self.b[l]), self.cout[l],
                      _lib.dptr(y), s)
            pres.append(y)
            if not last:
                a = torch.empty((B, H, W, 32), dtype=torch.int16, device="cuda")
                _lib.call("fsx_cnn_act", _lib.dptr(y), npix, 32, 32, 1, _lib.dptr(a), s)
                acts.append(a)
        # loss gradient w.r.t. the output, BF16 with 8 channels
        d = torch.empty((B, H, W, 8), dtype=torch.int16, device="cuda")
        part = torch.empty(256, dtype=torch.float64, device="cuda")
        loss = ctypes.c_double(0.0)
        tc = target.contiguous()
        _lib.call("fsx_cnn_mse", _lib.dptr(pres[-1]), _lib.dptr(tc), npix, 8,
                  _lib.dptr(d), _lib.dptr(part), ctypes.byref(loss) if want_loss else None, s)
        grads_w = [None] * self.nlayers
        grads_b = [None] * self.nlayers
        nbytes = ctypes.c_longlong(0)
        for l in range(self.nlayers - 1, -1, -1):
            ci, co_pad = self.cin[l], (8 if l == self.nlayers - 1 else 32)
            _lib.call("fsx_cnn_wgrad_workspace", B, H, ci, ctypes.byref(nbytes))
            work = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
            gw = torch.empty((32, 25, ci), **dev)
            gb = torch.empty(32, **dev)
            _lib.call("fsx_cnn_conv_wgrad", _lib.dptr(acts[l]), ci, _lib.dptr(d), co_pad, B, H,
                      W, _lib.dptr(gw), _lib.dptr(gb), _lib.dptr(work), nbytes.value, s)
            grads_w[l], grads_b[l] = gw, gb
            if l == 0:
                break
            # input gradient: flipped, transposed weights on d
            wt = torch.empty((32, 25, co_pad), dtype=torch.int16, device="cuda")
            _lib.call("fsx_cnn_conv_flip", _lib.dptr(self.wb[l]), ci, self.cout[l], co_pad,
                      _lib.dptr(wt), s)
            g = torch.empty((B, H, W, 32), **dev)
            _lib.call("fsx_cnn_conv_fwd", _lib.dptr(d), B, H, W, co_pad, _lib.dptr(wt),
                      _lib.dptr(self.zero_bias), 32, _lib.dptr(g), s)
            d = torch.empty((B, H, W, 32), dtype=torch.int16, device="cuda")
            _lib.call("fsx_cnn_act_bwd", _lib.dptr(g), 32, _lib.dptr(pres[l - 1]), npix, 32,
                      _lib.dptr(d), s)
        self.step_count += 1
        lr = self.lr if lr is None else lr
        for i, (p, gr) in enumerate(zip(self.w + self.b, grads_w + grads_b)):
            wb = self.wb[i] if i < self.nlayers else None
            _lib.call("fsx_cnn_adam", _lib.dptr(p), _lib.dptr(gr), _lib.dptr(self.m1[i]),
                      _lib.dptr(self.m2[i]), p.numel(), float(lr), 0.9, 0.999, 1e-8,
                      self.step_count, _lib.dptr(wb) if wb is not None else None, s)
        self.last_grads = (grads_w, grads_b)
        return loss.value if want_loss else None
