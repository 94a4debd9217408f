"""fsx_cnn_conv_wgrad_noalloc against fsx_cnn_conv_wgrad and the oracle.

The entry point without allocation shares the weight-gradient kernels and
their partials layout; its partials live in a static device array of the
library.  Bar: bit-identical to fsx_cnn_conv_wgrad on the same inputs, and
within 1e-5 of the largest value of the FP64 oracle (oracle/cnn.py
conv_weight_grad), on two calls in a row, the second with a smaller input;
the device's free memory does not change over a call.
"""

import ctypes

import numpy as np
import pytest

from fluidsim_amd import _lib
from fluidsim_amd.closure import bf16_bits, bf16_value
from oracle import cnn as C

pytestmark = pytest.mark.gpu


def _rounded(a):
    return bf16_value(bf16_bits(a)).astype(np.float64)


def _relmax(got, ref):
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


def _case(rng, ci, co_in, shape):
    import torch

    x = _rounded(rng.standard_normal(shape + (ci,)))
    dy = _rounded(rng.standard_normal(shape + (co_in,)))
    xd = torch.from_numpy(bf16_bits(x)).cuda()
    dyd = torch.from_numpy(bf16_bits(dy)).cuda()
    return x, dy, xd, dyd


def _outputs(ci):
    import torch

    return (torch.empty((32, 25, ci), dtype=torch.float32, device="cuda"),
            torch.empty(32, dtype=torch.float32, device="cuda"))


def test_noalloc_weight_gradient_two_calls():
    import torch

    rng = np.random.default_rng(11)
    s = _lib.stream_ptr()
    cases = [(32, 32, (8, 128, 128)), (8, 32, (2, 8, 32))]  # second call smaller
    prepared = []
    for ci, co_in, shape in cases:
        x, dy, xd, dyd = _case(rng, ci, co_in, shape)
        nb = ctypes.c_longlong(0)
        _lib.call("fsx_cnn_wgrad_workspace", shape[0], shape[1], ci, ctypes.byref(nb))
        work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
        prepared.append((ci, co_in, shape, x, dy, xd, dyd, work, nb.value,
                         _outputs(ci), _outputs(ci)))
    torch.cuda.synchronize()
    for ci, co_in, shape, x, dy, xd, dyd, work, nbytes, (dw1, db1), (dw2, db2) in prepared:
        B, H, W = shape
        print(f"case no-alloc wgrad ci={ci} co_in={co_in} shape={shape}", flush=True)
        _lib.call("fsx_cnn_conv_wgrad", _lib.dptr(xd), ci, _lib.dptr(dyd), co_in, B, H, W,
                  _lib.dptr(dw1), _lib.dptr(db1), _lib.dptr(work), nbytes, s)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        _lib.call("fsx_cnn_conv_wgrad_noalloc", _lib.dptr(xd), ci, _lib.dptr(dyd), co_in, B, H,
                  W, _lib.dptr(dw2), _lib.dptr(db2), s)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, (free0, free1)
        assert torch.equal(dw1, dw2) and torch.equal(db1, db2)
        rdw, rdb = C.conv_weight_grad(x, dy)
        got = dw2.cpu().numpy()
        assert _relmax(got[:co_in], rdw) < 1e-5
        assert np.all(got[co_in:] == 0)
        assert _relmax(db2.cpu().numpy()[:co_in], rdb) < 1e-5
