"""Argument errors of the C-ABI entry points (host logic; every check below
returns before any HIP call, so the tests run without a GPU).

Each call must return FSX_ERR_ARG (ValueError through the binding, as the
reference raises ValueError on bad input) with a message naming the
problem; no pointer passed here is ever dereferenced.
"""

import ctypes
import os

import pytest

from fluidsim_amd import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH),
                                reason="library not built")

# non-null stand-ins: the checks under test reject the call before any
# pointer is used
P = ctypes.c_void_p(0x1000)


def ints(*v):
    return (ctypes.c_int * len(v))(*v)


def ptrs(n):
    return (ctypes.c_void_p * n)(*([0x1000] * n))


CASES = [
    # convolution: channels, outputs, tile shape, fused-output buffers
    ("fsx_cnn_conv_fwd", (P, 1, 8, 32, 16, P, P, 32, P, None), "channels must be 8 or 32"),
    ("fsx_cnn_conv_fwd", (P, 1, 8, 32, 32, P, P, 33, P, None), "output channels"),
    ("fsx_cnn_conv_fwd", (P, 1, 6, 32, 32, P, P, 32, P, None), "multiple of 4"),
    ("fsx_cnn_conv_fwd", (P, 1, 8, 48, 32, P, P, 32, P, None), "multiple of 32"),
    ("fsx_cnn_conv_fwd", (None, 1, 8, 32, 32, P, P, 32, P, None), "bad argument"),
    ("fsx_cnn_conv_fwd", (P, 0, 8, 32, 32, P, P, 32, P, None), "bad argument"),
    ("fsx_cnn_conv_fwd_act", (P, 1, 8, 32, 32, P, P, P, None, None), "bad argument"),
    ("fsx_cnn_conv_fwd_act", (P, 1, 8, 32, 16, P, P, P, P, None), "channels must be 8 or 32"),
    ("fsx_cnn_conv_flip", (P, 32, 2, 1, P, None), "bad argument"),
    ("fsx_cnn_conv_flip", (P, 32, 1, 16, P, None), "bad argument"),
    # weight gradient and its workspace
    ("fsx_cnn_wgrad_workspace", (0, 8, 32, ctypes.byref(ctypes.c_longlong())), "bad argument"),
    ("fsx_cnn_wgrad_workspace", (1, 8, 16, ctypes.byref(ctypes.c_longlong())), "bad argument"),
    ("fsx_cnn_conv_wgrad", (P, 16, P, 32, 1, 8, 32, P, P, P, 1 << 20, None), "channels"),
    ("fsx_cnn_conv_wgrad", (P, 32, P, 32, 1, 6, 32, P, P, P, 1 << 20, None), "multiple of 4"),
    ("fsx_plan_set_tiles", (None, 0, 0, 0), "bad argument"),
    ("fsx_cnn_conv_wgrad_noalloc", (P, 16, P, 32, 1, 8, 32, P, P, None), "channels"),
    ("fsx_cnn_conv_wgrad_noalloc", (P, 32, P, 32, 1, 6, 32, P, P, None), "multiple of 4"),
    ("fsx_cnn_conv_wgrad_noalloc", (P, 32, P, 32, 1, 8, 36, P, P, None), "multiple of 4"),
    ("fsx_cnn_conv_wgrad_noalloc", (None, 32, P, 32, 1, 8, 32, P, P, None), "bad argument"),
    # elementwise pieces
    ("fsx_cnn_act", (P, 16, 4, 2, 1, P, None), "bad argument"),
    ("fsx_cnn_act_bwd", (P, 8, P, 16, 32, P, None), "bad argument"),
    ("fsx_cnn_mse", (P, P, 0, 8, P, P, None, None), "bad argument"),
    ("fsx_cnn_adam", (P, P, P, P, 10, 1e-3, 0.9, 0.999, 1e-8, 0, None, None), "bad argument"),
    ("fsx_cnn_adam_multi", (17, ptrs(17), ptrs(17), ptrs(17), ptrs(17), ints(*[8] * 17), None,
                            1e-3, 0.9, 0.999, 1e-8, 1, None), "1 to 16 tensors"),
    ("fsx_cnn_adam_multi", (2, ptrs(2), ptrs(2), ptrs(2), ptrs(2), ints(8, 0), None, 1e-3, 0.9,
                            0.999, 1e-8, 1, None), "bad argument"),
    ("fsx_cnn_to_bf16", (P, 0, P, None), "bad argument"),
    # network in one call: sizes, hidden widths, workspace
    ("fsx_cnn_forward", (0, ints(8), ints(1), ptrs(1), ptrs(1), P, 1, 8, 32, P, P, 1 << 20,
                         None), "bad argument"),
    ("fsx_cnn_forward", (2, ints(8, 32), ints(16, 1), ptrs(2), ptrs(2), P, 1, 8, 32, P, P,
                         1 << 20, None), "hidden layers have 32 channels"),
    ("fsx_cnn_forward", (2, ints(8, 32), ints(32, 1), ptrs(2), ptrs(2), P, 1, 8, 32, P, P,
                         100, None), "workspace too small"),
    ("fsx_cnn_forward", (2, ints(16, 32), ints(32, 1), ptrs(2), ptrs(2), P, 1, 8, 32, P, P,
                         1 << 20, None), "channels must be 8 or 32"),
    # closure term, timing, batches and graphs without a plan
    ("fsx_ns2d_closure_workspace", (None, ctypes.byref(ctypes.c_longlong())), "bad argument"),
    ("fsx_ns2d_closure_term", (None, 2, ints(8, 32), ints(32, 1), ptrs(2), ptrs(2), P, P, P,
                               1 << 20, None), "bad argument"),
    ("fsx_timing_read_network", (None, ctypes.byref(ctypes.c_double()),
                                 ctypes.byref(ctypes.c_double()),
                                 ctypes.byref(ctypes.c_longlong()), 0), "bad argument"),
    ("fsx_batch_set_lines", (None, 4), "not a batch plan"),
    ("fsx_graph_begin", (None,), "created stream"),
    ("fsx_graph_end", (None, None), "bad argument"),
    ("fsx_diag_layout", (None, 0, 8, None), "bad argument"),
    ("fsx_if_combine", (None, 1, P, 1, ptrs(1), None, None, None, None, None, None, 0, None),
     "bad argument"),
]


@pytest.mark.parametrize("name,args,message", CASES,
                         ids=[f"{c[0]}-{i}" for i, c in enumerate(CASES)])
def test_argument_error(name, args, message):
    with pytest.raises(ValueError) as exc:
        _lib.call(name, *args)
    assert message in str(exc.value)
