"""GPU tests of the closure network kernels (fsx_cnn_*) against the FP64
oracle (oracle/cnn.py) on the same BF16-rounded inputs and weights.

Bars: the MFMA operand layout is exact on small integers; convolutions,
input and weight gradients (FP32 accumulation of <= 800-term BF16
products) within 1e-5 of the largest value; the network forward within
5e-3 relative L2 of an oracle that rounds its activations to BF16 like the
kernels; the gradients of one training step (which pass through BF16)
within 3e-2 relative L2 of the FP64 gradients; training lowers the loss.
"""

import numpy as np
import pytest
import torch

from oracle import cnn as C
from fluidsim_amd import _lib
from fluidsim_amd.closure import CNNClosure, bf16_bits, bf16_value

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to("cuda") if dtype is None else t.to("cuda", dtype)


def rounded(a):
    return bf16_value(bf16_bits(a)).astype(np.float64)


def relmax(got, ref):
    return np.abs(got - ref).max() / np.abs(ref).max()


def test_mfma_layout_probe():
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, (32, 8)).astype(np.float32)
    B = rng.integers(-4, 5, (8, 32)).astype(np.float32)
    D = torch.zeros((32, 32), dtype=torch.float32, device="cuda")
    # device buffers are held in locals: a temporary freed after dptr() would
    # hand its memory to the next allocation before the kernel runs
    a, b = dev(bf16_bits(A)), dev(bf16_bits(B))
    _lib.call("fsx_cnn_probe", _lib.dptr(a), _lib.dptr(b), _lib.dptr(D), _lib.stream_ptr())
    assert np.array_equal(D.cpu().numpy(), A @ B)


def test_mfma16_layout_probe():
    """Operand layout of the K = 16 instruction the convolutions use."""
    rng = np.random.default_rng(11)
    A = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    B = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    D = torch.zeros((32, 32), dtype=torch.float32, device="cuda")
    a, b = dev(bf16_bits(A)), dev(bf16_bits(B))
    _lib.call("fsx_cnn_probe16", _lib.dptr(a), _lib.dptr(b), _lib.dptr(D), _lib.stream_ptr())
    assert np.array_equal(D.cpu().numpy(), A @ B)


def conv_gpu(x, w, b, co_out):
    B_, H, W, ci = x.shape
    y = torch.empty((B_, H, W, co_out), dtype=torch.float32, device="cuda")
    xd, wd, bd = dev(bf16_bits(x)), dev(bf16_bits(w)), dev(b.astype(np.float32))
    _lib.call("fsx_cnn_conv_fwd", _lib.dptr(xd), B_, H, W, ci, _lib.dptr(wd), _lib.dptr(bd),
              co_out, _lib.dptr(y), _lib.stream_ptr())
    return y.cpu().numpy()


@pytest.mark.parametrize("ci,shape", [(8, (2, 8, 32)), (32, (2, 8, 32)), (32, (1, 16, 128)),
                                      (8, (3, 12, 64)),
                                      # more 4-row blocks than two per CU: the
                                      # persistent form with its ring of patch rows
                                      (32, (20, 128, 32)), (8, (40, 64, 32))])
def test_conv_forward(ci, shape):
    rng = np.random.default_rng(1)
    x = rounded(rng.standard_normal(shape + (ci,)))
    w = rounded(rng.standard_normal((32, 25, ci)) * 0.2)
    b = rng.standard_normal(32)
    for co in (32, 8, 1):
        got = conv_gpu(x, w, b, co)
        ref = C.conv(x, w, b.astype(np.float32).astype(np.float64), co)
        print(f"conv ci={ci} {shape} co={co}: {relmax(got, ref):.2e}")
        assert relmax(got, ref) < 1e-5


@pytest.mark.parametrize("ci,co,co_pad", [(32, 32, 32), (32, 1, 8), (8, 32, 32)])
def test_input_gradient_by_flip(ci, co, co_pad):
    rng = np.random.default_rng(2)
    B_, H, W = 2, 8, 32
    w = np.zeros((32, 25, ci))
    w[:co] = rounded(rng.standard_normal((co, 25, ci)) * 0.2)
    dy = np.zeros((B_, H, W, co_pad))
    dy[..., :co] = rounded(rng.standard_normal((B_, H, W, co)))
    wt = torch.empty((32, 25, co_pad), dtype=torch.int16, device="cuda")
    s = _lib.stream_ptr()
    wd = dev(bf16_bits(w))
    _lib.call("fsx_cnn_conv_flip", _lib.dptr(wd), ci, co, co_pad, _lib.dptr(wt), s)
    g = torch.empty((B_, H, W, ci), dtype=torch.float32, device="cuda")
    zb = torch.zeros(32, dtype=torch.float32, device="cuda")
    dyd = dev(bf16_bits(dy))
    _lib.call("fsx_cnn_conv_fwd", _lib.dptr(dyd), B_, H, W, co_pad, _lib.dptr(wt),
              _lib.dptr(zb), ci, _lib.dptr(g), s)
    ref = C.conv_input_grad(dy[..., :co], w, ci)
    assert relmax(g.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("ci,co_in,shape", [(8, 32, (2, 8, 32)), (32, 32, (2, 8, 32)),
                                            (32, 8, (3, 16, 64)), (8, 8, (1, 4, 16)),
                                            (32, 32, (2, 4, 24)), (8, 32, (1, 8, 40)),
                                            (32, 32, (8, 128, 128)),
                                            # more row blocks than workgroups: each
                                            # workgroup keeps its x rows in the ring
                                            (32, 32, (16, 128, 64)), (8, 32, (20, 64, 32)),
                                            (32, 8, (9, 128, 40)),
                                            # too wide for 4-row blocks: 2-row blocks
                                            (32, 32, (1, 8, 256)), (8, 32, (2, 12, 248))])
def test_weight_gradient(ci, co_in, shape):
    rng = np.random.default_rng(3)
    x = rounded(rng.standard_normal(shape + (ci,)))
    dy = rounded(rng.standard_normal(shape + (co_in,)))
    B_, H, W = shape
    nb = ctypes_ll()
    _lib.call("fsx_cnn_wgrad_workspace", B_, H, ci, nb)
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    dw = torch.empty((32, 25, ci), dtype=torch.float32, device="cuda")
    db = torch.empty(32, dtype=torch.float32, device="cuda")
    xd, dyd = dev(bf16_bits(x)), dev(bf16_bits(dy))
    _lib.call("fsx_cnn_conv_wgrad", _lib.dptr(xd), ci, _lib.dptr(dyd), co_in, B_, H, W,
              _lib.dptr(dw), _lib.dptr(db), _lib.dptr(work), nb.value, _lib.stream_ptr())
    rdw, rdb = C.conv_weight_grad(x, dy)
    got = dw.cpu().numpy()
    assert relmax(got[:co_in], rdw) < 1e-5
    assert np.all(got[co_in:] == 0)
    assert relmax(db.cpu().numpy()[:co_in], rdb) < 1e-5


def ctypes_ll():
    import ctypes

    return ctypes.c_longlong(0)


def oracle_net(net):
    layers = []
    for l in range(net.nlayers):
        w = rounded(net.w[l].cpu().numpy())
        co = net.cout[l]
        layers.append((w[:co], net.b[l].cpu().numpy()[:co].astype(np.float64)))
    return C.Net(layers)


def forward_bf16_acts(onet, x):
    """Oracle forward rounding the activations to BF16 like the kernels."""
    a = rounded(x)
    for i, (w, b) in enumerate(onet.layers):
        y = C.conv(a, w[:, :, : a.shape[3]], b)
        a = rounded(C.gelu(y)) if i < len(onet.layers) - 1 else y
    return a[..., 0]


def test_network_forward():
    rng = np.random.default_rng(4)
    net = CNNClosure(n_in=1, nlayers=6, seed=5)
    x = rng.standard_normal((2, 16, 64, 1)).astype(np.float32)
    got = net.forward(dev(x)).cpu().numpy()
    onet = oracle_net(net)
    ref = forward_bf16_acts(onet, np.concatenate([x, np.zeros(x.shape[:3] + (7,))], axis=3))
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"network forward rel L2 = {err:.2e}")
    assert err < 5e-3


def test_training_step_gradients_and_descent():
    rng = np.random.default_rng(6)
    net = CNNClosure(n_in=1, nlayers=4, seed=7, lr=2e-3)
    B_, H, W = 4, 16, 32
    x = rng.standard_normal((B_, H, W, 1)).astype(np.float32)
    # a smooth, learnable target: a local average of the input
    xr = x[..., 0].astype(np.float64)
    target = sum(np.roll(xr, (dy, dx), axis=(1, 2)) for dy in (-1, 0, 1) for dx in (-1, 0, 1)) / 9
    onet = oracle_net(net)
    xpad = np.concatenate([rounded(x), np.zeros(x.shape[:3] + (7,))], axis=3)
    loss_ref, grads_ref, _ = onet.gradients(xpad, target)
    loss = net.train_step(dev(x), dev(target.astype(np.float32)))
    assert loss == pytest.approx(loss_ref, rel=5e-3)
    gw, gb = net.last_grads
    for l in range(net.nlayers):
        co = net.cout[l]
        got = gw[l].cpu().numpy()[:co, :, : grads_ref[l][0].shape[2]]
        err = np.linalg.norm(got - grads_ref[l][0]) / np.linalg.norm(grads_ref[l][0])
        errb = np.linalg.norm(gb[l].cpu().numpy()[:co] - grads_ref[l][1]) / np.linalg.norm(
            grads_ref[l][1])
        print(f"layer {l}: dw rel L2 = {err:.2e}, db rel L2 = {errb:.2e}")
        assert err < 3e-2 and errb < 3e-2
    losses = [net.train_step(dev(x), dev(target.astype(np.float32))) for _ in range(200)]
    print(f"loss {loss:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.1 * loss


@pytest.mark.parametrize("ci", [8, 32])
@pytest.mark.parametrize("shape", [(2, 8, 64), (20, 128, 32)])
def test_fused_activation_epilogue_is_bit_exact(ci, shape):
    """fsx_cnn_conv_fwd_act (GELU in the convolution epilogue) writes the
    same pre-activations and BF16 activations as fsx_cnn_conv_fwd followed
    by fsx_cnn_act (also in the persistent form of large batches)."""
    rng = np.random.default_rng(8)
    B_, H, W = shape
    x = dev(bf16_bits(rng.standard_normal((B_, H, W, ci))))
    w = dev(bf16_bits(rng.standard_normal((32, 25, ci)) * 0.2))
    b = dev(rng.standard_normal(32).astype(np.float32))
    s = _lib.stream_ptr()
    y0 = torch.empty((B_, H, W, 32), dtype=torch.float32, device="cuda")
    a0 = torch.empty((B_, H, W, 32), dtype=torch.int16, device="cuda")
    _lib.call("fsx_cnn_conv_fwd", _lib.dptr(x), B_, H, W, ci, _lib.dptr(w), _lib.dptr(b), 32,
              _lib.dptr(y0), s)
    _lib.call("fsx_cnn_act", _lib.dptr(y0), B_ * H * W, 32, 32, 1, _lib.dptr(a0), s)
    y1 = torch.full_like(y0, np.nan)
    a1 = torch.zeros_like(a0)
    a2 = torch.zeros_like(a0)
    _lib.call("fsx_cnn_conv_fwd_act", _lib.dptr(x), B_, H, W, ci, _lib.dptr(w), _lib.dptr(b),
              _lib.dptr(y1), _lib.dptr(a1), s)
    _lib.call("fsx_cnn_conv_fwd_act", _lib.dptr(x), B_, H, W, ci, _lib.dptr(w), _lib.dptr(b),
              None, _lib.dptr(a2), s)
    assert torch.equal(y0, y1)
    assert torch.equal(a0, a1) and torch.equal(a0, a2)


@pytest.mark.parametrize("ci", [8, 32])
@pytest.mark.parametrize("shape", [(2, 8, 64), (20, 128, 32)])
def test_fused_input_gradient_epilogue_is_bit_exact(ci, shape):
    """fsx_cnn_conv_bwd_act (the input gradient taken through the previous
    layer's GELU in the convolution epilogue) writes the same BF16 gradient
    as fsx_cnn_conv_fwd into FP32 followed by fsx_cnn_act_bwd (also in the
    persistent form of large batches), with a zero and a nonzero bias."""
    rng = np.random.default_rng(11)
    B_, H, W = shape
    x = dev(bf16_bits(rng.standard_normal((B_, H, W, ci))))
    w = dev(bf16_bits(rng.standard_normal((32, 25, ci)) * 0.2))
    pre = dev(rng.standard_normal((B_, H, W, 32)).astype(np.float32) * 2)
    s = _lib.stream_ptr()
    for b in (torch.zeros(32, device="cuda"), dev(rng.standard_normal(32).astype(np.float32))):
        g = torch.empty((B_, H, W, 32), dtype=torch.float32, device="cuda")
        d0 = torch.empty((B_, H, W, 32), dtype=torch.int16, device="cuda")
        _lib.call("fsx_cnn_conv_fwd", _lib.dptr(x), B_, H, W, ci, _lib.dptr(w), _lib.dptr(b),
                  32, _lib.dptr(g), s)
        _lib.call("fsx_cnn_act_bwd", _lib.dptr(g), 32, _lib.dptr(pre), B_ * H * W, 32,
                  _lib.dptr(d0), s)
        d1 = torch.zeros_like(d0)
        _lib.call("fsx_cnn_conv_bwd_act", _lib.dptr(x), B_, H, W, ci, _lib.dptr(w), _lib.dptr(b),
                  _lib.dptr(pre), _lib.dptr(d1), s)
        assert torch.equal(d0, d1)
    with pytest.raises(ValueError):
        _lib.call("fsx_cnn_conv_bwd_act", _lib.dptr(x), B_, H, W, ci, _lib.dptr(w),
                  _lib.dptr(b), None, _lib.dptr(d1), s)


def test_activations_match_the_oracle_to_bf16():
    """fsx_cnn_act (GELU) and fsx_cnn_act_bwd (times GELU') against the FP64
    oracle: within one BF16 unit in the last place of the rounded oracle
    value everywhere (the device erf is within 1.5e-7 of erf), and equal to
    it for nearly all elements at y >= -1; inputs cover the polynomial's whole range,
    zeros, and large magnitudes where GELU is 0 or the identity."""
    rng = np.random.default_rng(13)
    n = 1 << 16
    y = np.concatenate([rng.standard_normal(n - 8) * 3, [0.0, -0.0, 1e-6, -1e-6, 12, -12, 40, -40]])
    y = y.astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    s = _lib.stream_ptr()
    yd, gd = dev(y), dev(g)  # held: the launches only take their addresses
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    _lib.call("fsx_cnn_act", _lib.dptr(yd), n, 1, 1, 1, _lib.dptr(a), s)
    d = torch.empty(n, dtype=torch.int16, device="cuda")
    _lib.call("fsx_cnn_act_bwd", _lib.dptr(gd), 1, _lib.dptr(yd), n, 1, _lib.dptr(d), s)
    y64, g64 = y.astype(np.float64), g.astype(np.float64)
    # 1 + erf(x) cancels in FP32 for large negative x (as with erff): an
    # absolute allowance of 2^-22 |y| (times |g|) besides the BF16 unit
    for got_bits, ref, scale in ((a, C.gelu(y64), np.abs(y64)),
                                 (d, g64 * C.gelu_grad(y64), np.abs(g64 * y64))):
        got = bf16_value(got_bits.cpu().numpy().view(np.uint16)).astype(np.float64)
        r = rounded(ref)
        tol = np.abs(r) * 2.0 ** -7 + scale * 2.0 ** -22
        assert np.all(np.abs(got - r) <= tol), np.abs(got - r).max()
        # equal where 1 + erf does not cancel
        assert np.mean(got[y64 >= -1] == r[y64 >= -1]) > 0.999


def test_adam_multi_tensor_is_bit_exact():
    """fsx_cnn_adam_multi (all tensors in one launch) updates weights, both
    moments and the BF16 copies exactly as one fsx_cnn_adam call per tensor,
    over two steps, with tensors with and without a BF16 copy and sizes
    above one grid's worth of workgroups."""
    import ctypes

    rng = np.random.default_rng(12)
    sizes = [32 * 25 * 8, 32, 32 * 25 * 32, 1 << 21, 7]
    s = _lib.stream_ptr()

    def state():
        return [[dev(rng.standard_normal(n).astype(np.float32)) for n in sizes] for _ in range(3)]

    w0, m10, m20 = state()
    m10 = [t.abs() * 0.1 for t in m10]
    m20 = [t.abs() * 0.01 for t in m20]
    one = [[t.clone() for t in ts] for ts in (w0, m10, m20)]
    many = [[t.clone() for t in ts] for ts in (w0, m10, m20)]
    wb1 = [torch.zeros(n, dtype=torch.int16, device="cuda") if i % 2 == 0 else None
           for i, n in enumerate(sizes)]
    wb2 = [t.clone() if t is not None else None for t in wb1]
    nt = len(sizes)

    def ptrs(ts):
        return (ctypes.c_void_p * nt)(*[t.data_ptr() if t is not None else None for t in ts])

    for step in (1, 2):
        g = [dev(rng.standard_normal(n).astype(np.float32)) for n in sizes]
        for i in range(nt):
            _lib.call("fsx_cnn_adam", _lib.dptr(one[0][i]), _lib.dptr(g[i]), _lib.dptr(one[1][i]),
                      _lib.dptr(one[2][i]), sizes[i], 1e-3, 0.9, 0.999, 1e-8, step,
                      _lib.dptr(wb1[i]) if wb1[i] is not None else None, s)
        _lib.call("fsx_cnn_adam_multi", nt, ptrs(many[0]), ptrs(g), ptrs(many[1]),
                  ptrs(many[2]), (ctypes.c_int * nt)(*sizes), ptrs(wb2), 1e-3, 0.9, 0.999, 1e-8,
                  step, s)
    for a, b in zip(sum(one, []), sum(many, [])):
        assert torch.equal(a, b)
    for a, b in zip(wb1, wb2):
        assert a is None or torch.equal(a, b)
    assert not torch.equal(one[0][0], w0[0])


def test_network_forward_matches_layer_by_layer():
    """fsx_cnn_forward (fused hidden layers) equals the training path's
    layer-by-layer forward bit for bit."""
    rng = np.random.default_rng(9)
    net = CNNClosure(n_in=1, nlayers=5, seed=2)
    for t in net.b:
        t.copy_(torch.from_numpy(rng.standard_normal(32).astype(np.float32) * 0.1))
    x = dev(rng.standard_normal((2, 16, 64, 1)).astype(np.float32))
    got = net.forward(x)
    s = _lib.stream_ptr()
    a = net.pack_input(x).clone()
    for l in range(net.nlayers):
        y = torch.empty((2, 16, 64, net.cout[l]), dtype=torch.float32, device="cuda")
        _lib.call("fsx_cnn_conv_fwd", _lib.dptr(a), 2, 16, 64, net.cin[l], _lib.dptr(net.wb[l]),
                  _lib.dptr(net.b[l]), net.cout[l], _lib.dptr(y), s)
        if l < net.nlayers - 1:
            a = torch.empty((2, 16, 64, 32), dtype=torch.int16, device="cuda")
            _lib.call("fsx_cnn_act", _lib.dptr(y), 2 * 16 * 64, 32, 32, 1, _lib.dptr(a), s)
    assert torch.equal(got, y[..., 0])
