"""ctypes binding of the HIP engine library ``libfluidsim_amd.so``.

The library is built in-tree by ``fluidsim_amd.build.build()`` (hipcc,
gfx950).  There is no CPU fallback: if the library is missing, importing the
solvers works but any engine call raises ``RuntimeError``.
"""

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FLUIDSIM_AMD_LIB selects another build of the engine library (e.g. a
# variant compiled with different tile sizes).
LIB_PATH = os.environ.get("FLUIDSIM_AMD_LIB") or os.path.join(HERE, "libfluidsim_amd.so")

FSX_ERR_ARG = -1
FSX_ERR_HIP = -2

_lock = threading.Lock()
_lib = None

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P = ctypes.c_void_p

# name -> argtypes (all functions return int)
SIGNATURES = {
    "fsx_version": [],
    "fsx_last_error": [c_char_p, c_int],
    "fsx_plan_create": [
        ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int, c_double,
        c_double, c_double, c_int, P, c_int,
    ],
    "fsx_plan_destroy": [P],
    "fsx_plan_info": [P, ctypes.POINTER(c_int), c_int],
    "fsx_fft": [P, P, P, c_int, P],
    "fsx_ifft": [P, P, P, c_int, P],
    "fsx_ns2d_tendency": [P, P, P, c_double, c_int, P],
    "fsx_ns3d_tendency": [P, P, P, c_int, c_double, c_int, P],
    "fsx_rk_stage": [
        P, c_int, c_int, c_int, P, P, P, P, c_int, c_int,
        ctypes.POINTER(c_double), P,
    ],
    "fsx_ns2d_rk_stage": [
        P, c_int, c_int, P, P, P, c_double, ctypes.POINTER(c_double), c_double,
        c_double, P,
    ],
    "fsx_ns2d_rk_stage_ext": [
        P, c_int, c_int, P, P, P, c_double, ctypes.POINTER(c_double), c_double,
        c_double, P, P,
    ],
    "fsx_ns3d_rk_stage": [
        P, c_int, c_int, P, P, P, c_int, c_double, c_int, c_int,
        ctypes.POINTER(c_double), c_double, c_double, P,
    ],
    "fsx_strat_rk_stage": [
        P, c_int, c_int, P, P, P, c_int, c_double, c_double, c_int, c_int,
        ctypes.POINTER(c_double), c_double, c_double, P,
    ],
    "fsx_strat_tendency": [P, P, P, c_int, c_double, c_double, c_int, c_int, c_int, P],
    "fsx_project": [P, P, P, P, c_int, P],
    "fsx_dealias": [P, P, c_int, c_int, P],
    "fsx_set_dt": [P, c_double, P],
    "fsx_reset_maxima": [P, P],
    "fsx_cfl_update": [P, c_int, c_double, c_double, P],
    "fsx_get_scalars": [P, ctypes.POINTER(c_double), P],
    "fsx_clear_nan_flag": [P, P],
    "fsx_sums": [
        P, c_int, P, P, P, P, c_double, ctypes.POINTER(c_double),
        ctypes.POINTER(c_double), P,
    ],
    "fsx_forcing_setup": [
        P, c_int, c_int, c_int, c_int, c_int, P, P, c_int, c_int,
    ],
    "fsx_forcing_set_pair": [P, P, P, P],
    "fsx_forcing_set_coarse": [P, P, P],
    "fsx_forcing_enable": [P, c_int],
    "fsx_forcing_step": [P, c_double, c_double, c_int],
    "fsx_forcing_normalize": [P, P, P],
    "fsx_forcing_graph": [P, c_int],
    "fsx_forcing_push": [P, P],
    "fsx_forcing_add": [P, P, c_int, c_double, P],
    "fsx_strain_rate": [P, P, P, P],
    "fsx_smagorinsky_phys": [P, P, c_double, P],
    "fsx_smagorinsky_div": [P, P, P, c_int, P],
    "fsx_if_combine": [
        P, c_int, P, c_int, P, ctypes.POINTER(c_double), ctypes.POINTER(c_double),
        ctypes.POINTER(c_double), ctypes.POINTER(c_int), ctypes.POINTER(c_double),
        ctypes.POINTER(c_double), c_int, P,
    ],
    "fsx_resize_fft": [P, c_int, c_int, c_int, P, c_int, c_int, c_int, c_int, c_int, P],
    "fsx_diag_layout": [P, c_int, c_int, ctypes.POINTER(ctypes.c_longlong)],
    "fsx_diag": [P, c_int, c_int, P, c_int, ctypes.POINTER(c_double),
                 ctypes.POINTER(c_double), P],
    "fsx_ns2d_budget_fields": [P, P, c_double, P, P, P, P],
    "fsx_plan_create_batch": [P, c_int, c_int, c_double, c_double, c_int, P, c_int],
    "fsx_batch_set_params": [P, ctypes.POINTER(c_double), ctypes.POINTER(c_double)],
    "fsx_batch_set_row": [P, c_int],
    "fsx_plan_set_tiles": [P, c_int, c_int, c_int],
    "fsx_batch_set_lines": [P, c_int],
    "fsx_ns2d_batch_rk_stage": [P, c_int, c_int, P, P, P, c_double, P],
    "fsx_graph_begin": [P],
    "fsx_graph_end": [P, P],
    "fsx_graph_launch": [P, P],
    "fsx_graph_destroy": [P],
    "fsx_cnn_probe": [P, P, P, P],
    "fsx_cnn_probe16": [P, P, P, P],
    "fsx_cnn_conv_fwd": [P, c_int, c_int, c_int, c_int, P, P, c_int, P, P],
    "fsx_cnn_conv_fwd_act": [P, c_int, c_int, c_int, c_int, P, P, P, P, P],
    "fsx_cnn_conv_bwd_act": [P, c_int, c_int, c_int, c_int, P, P, P, P, P],
    "fsx_cnn_conv_flip": [P, c_int, c_int, c_int, P, P],
    "fsx_cnn_wgrad_workspace": [c_int, c_int, c_int, ctypes.POINTER(ctypes.c_longlong)],
    "fsx_cnn_conv_wgrad": [
        P, c_int, P, c_int, c_int, c_int, c_int, P, P, P, ctypes.c_longlong, P,
    ],
    "fsx_cnn_conv_wgrad_noalloc": [P, c_int, P, c_int, c_int, c_int, c_int, P, P, P],
    "fsx_cnn_act": [P, ctypes.c_longlong, c_int, c_int, c_int, P, P],
    "fsx_cnn_act_bwd": [P, c_int, P, ctypes.c_longlong, c_int, P, P],
    "fsx_cnn_mse": [P, P, ctypes.c_longlong, c_int, P, P, P, P],
    "fsx_cnn_adam": [
        P, P, P, P, c_int, c_double, c_double, c_double, c_double, c_int, P, P,
    ],
    "fsx_cnn_adam_multi": [
        c_int, P, P, P, P, ctypes.POINTER(c_int), P, c_double, c_double, c_double, c_double,
        c_int, P,
    ],
    "fsx_cnn_to_bf16": [P, c_int, P, P],
    "fsx_cnn_forward": [
        c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int), P, P, P, c_int, c_int, c_int,
        P, P, ctypes.c_longlong, P,
    ],
    "fsx_ns2d_closure_workspace": [P, ctypes.POINTER(ctypes.c_longlong)],
    "fsx_ns2d_closure_term": [
        P, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int), P, P, P, P, P,
        ctypes.c_longlong, P,
    ],
    "fsx_timing_enable": [P, c_int],
    "fsx_timing_read_network": [
        P, ctypes.POINTER(c_double), ctypes.POINTER(c_double),
        ctypes.POINTER(ctypes.c_longlong), c_int,
    ],
    "fsx_timing_read": [
        P, ctypes.POINTER(c_double), ctypes.POINTER(c_double),
        ctypes.POINTER(ctypes.c_longlong), c_int,
    ],
}


def load_library(path=LIB_PATH):
    """Load the shared library and declare the exported signatures."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"fluidsim_amd engine library not found at {path}; build it with "
            "`python -c 'import fluidsim_amd.build as b; b.build()'`"
        )
    # torch bundles its own libamdhip64 (SONAME libamdhip64.so.7).  Loading it
    # first makes the engine bind to that same runtime instead of pulling a
    # second HIP runtime from /opt/rocm into the process: device pointers and
    # streams are shared with torch, so there must be exactly one runtime.
    import torch  # noqa: F401

    lib = ctypes.CDLL(path)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = c_int
    return lib


def get_lib():
    global _lib
    with _lock:
        if _lib is None:
            _lib = load_library()
        return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    get_lib().fsx_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    """Map a return code to the exception type the reference raises."""
    if rc == 0:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == FSX_ERR_ARG:
        raise ValueError(msg)
    raise RuntimeError(msg)


def call(name, *args):
    rc = getattr(get_lib(), name)(*args)
    check(rc, name)


def dptr(t):
    """Device pointer of a torch tensor (must be contiguous)."""
    if not t.is_contiguous():
        raise ValueError("engine buffers must be contiguous")
    return c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def doubles(values):
    arr = (c_double * len(values))(*values)
    return arr


class Plan:
    """Owner of a C-ABI plan (operator grid, FFT tables, workspace)."""

    def __init__(self, dim, nx, ny, nz, Lx, Ly, Lz, precision, where_dealiased,
                 nf_max):
        self.lib = get_lib()
        self._h = c_void_p()
        mask_ptr = None
        if where_dealiased is not None:
            self._mask = np.ascontiguousarray(where_dealiased, dtype=np.uint8)
            mask_ptr = self._mask.ctypes.data_as(c_void_p)
        rc = self.lib.fsx_plan_create(
            ctypes.byref(self._h), dim, nx, ny, nz, float(Lx), float(Ly),
            float(Lz), precision, mask_ptr, nf_max,
        )
        check(rc, "fsx_plan_create")
        info = (c_int * 12)()
        check(self.lib.fsx_plan_info(self._h, info, 12))
        self.info = list(info)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.fsx_plan_destroy(h)
            except Exception:
                pass
            self._h = None
