"""Helpers shared by the GPU tests: device buffers and direct C-ABI calls."""

import numpy as np

from fluidsim_amd import _lib


def torch():
    import torch as _t

    return _t


def to_dev_complex(a, prec):
    t = torch()
    ct = t.complex64 if prec == 4 else t.complex128
    return t.from_numpy(np.ascontiguousarray(a)).to(ct).to("cuda").contiguous()


def to_dev_real(a, prec):
    t = torch()
    rt = t.float32 if prec == 4 else t.float64
    return t.from_numpy(np.ascontiguousarray(a)).to(rt).to("cuda").contiguous()


def zeros_complex(shape, prec):
    t = torch()
    ct = t.complex64 if prec == 4 else t.complex128
    return t.zeros(shape, dtype=ct, device="cuda")


def zeros_real(shape, prec):
    t = torch()
    rt = t.float32 if prec == 4 else t.float64
    return t.zeros(shape, dtype=rt, device="cuda")


def host(t):
    torch().cuda.synchronize()
    return t.cpu().numpy()


def make_plan(oper, prec, nf_max, dealias=True):
    """C-ABI plan for an oracle operator (same grid and mask)."""
    mask = oper.where_dealiased if dealias else None
    if len(oper.shapeX) == 2:
        return _lib.Plan(2, oper.nx, oper.ny, 1, oper.Lx, oper.Ly, 1.0, prec,
                         mask, nf_max)
    return _lib.Plan(3, oper.nx, oper.ny, oper.nz, oper.Lx, oper.Ly, oper.Lz,
                     prec, mask, nf_max)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) /
                 max(np.linalg.norm(np.asarray(b)), 1e-300))
