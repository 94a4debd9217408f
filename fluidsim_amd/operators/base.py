"""Device pseudo-spectral operator base shared by the 2D and 3D operators.

Transforms, dealiasing and projections run in the HIP engine through the
C-ABI plan; the small arithmetic helpers (curl, gradient, ...) are torch
elementwise expressions on device K arrays.  Methods accept numpy arrays
(copied to the GPU, result returned as numpy, like fluidfft's operators) or
torch CUDA tensors (result stays on the device).
"""

import numpy as np

from .. import _lib
from ..setofvariables import SetOfVariables


def _torch():
    import torch

    return torch


PRECISIONS = {"float64": 8, "double": 8, "float32": 4, "single": 4}


def k_signed(n, dk):
    idx = np.arange(n)
    return dk * np.where(idx <= n // 2, idx, idx - n).astype(np.float64)


class OperatorsBase:
    """Common machinery; subclasses define the grid and _kvectors()."""

    dim = None

    def _init_precision(self, params):
        precision = getattr(params.oper, "precision", "float64")
        if precision not in PRECISIONS:
            raise ValueError(
                f"params.oper.precision must be one of {sorted(PRECISIONS)}"
            )
        self.precision = PRECISIONS[precision]
        torch = _torch()
        if self.precision == 8:
            self.dtype_real, self.dtype_complex = torch.float64, torch.complex128
            self.np_real, self.np_complex = np.float64, np.complex128
        else:
            self.dtype_real, self.dtype_complex = torch.float32, torch.complex64
            self.np_real, self.np_complex = np.float32, np.complex64

    def _init_plan(self, nf_max):
        mask = self.where_dealiased if self._has_to_dealiase else None
        nz = getattr(self, "nz", 1)
        Lz = getattr(self, "Lz", 1.0)
        self.plan = _lib.Plan(
            self.dim, self.nx, self.ny, nz, self.Lx, self.Ly, Lz,
            self.precision, mask, nf_max,
        )

    # ------------------------------------------------------------ masks
    def _compute_where_dealiased(self, truncation_shape):
        """cubic: |k_i| >= c dk_i n_i / 2 along any axis (fluidfft rule);
        spherical / no_multiple_aliases: fluidsim/operators/base.py:47-72."""
        c = self.coef_dealiasing
        shapeK = self.shapeK
        if not c:
            return np.zeros(shapeK, dtype=np.uint8)
        ks = self._axes_k_1d()
        cubic = np.zeros(shapeK, dtype=bool)
        for axis, (k1d, dk, n) in enumerate(ks):
            shp = [1] * len(shapeK)
            shp[axis] = len(k1d)
            cubic |= (np.abs(k1d) >= c * dk * n / 2).reshape(shp)
        if truncation_shape == "cubic":
            return np.array(cubic, dtype=np.uint8)
        K2 = self._K2_host()
        kmax = c * self.deltakx * self.nx / 2
        if truncation_shape == "spherical":
            return np.array(K2 >= kmax**2, dtype=np.uint8)
        if truncation_shape == "no_multiple_aliases":
            al = []
            for axis, (k1d, dk, n) in enumerate(ks):
                shp = [1] * len(shapeK)
                shp[axis] = len(k1d)
                al.append(
                    np.broadcast_to(
                        (np.abs(k1d) >= 2 / 3 * dk * n / 2).reshape(shp), shapeK
                    )
                )
            if len(al) == 2:
                m = al[0] & al[1]
            else:
                m = (al[0] & al[1]) | (al[1] & al[2]) | (al[2] & al[0])
            return np.array(m | (K2 >= kmax**2), dtype=np.uint8)
        raise ValueError(
            'truncation_shape must be "cubic", "spherical" or "no_multiple_aliases"'
        )

    # ------------------------------------------------------------ arrays
    def _is_np(self, a):
        return isinstance(a, np.ndarray) or np.isscalar(a)

    def _to_dev(self, a, complex_):
        torch = _torch()
        dt = self.dtype_complex if complex_ else self.dtype_real
        if isinstance(a, torch.Tensor):
            if a.device.type == "cuda" and a.dtype == dt and a.is_contiguous():
                return a
            return a.to(device="cuda", dtype=dt).contiguous()
        return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dt)

    def create_arrayX(self, value=None, shape=None):
        torch = _torch()
        shp = tuple(self.shapeX) if shape is None else tuple(shape)
        if value is None:
            return torch.empty(shp, dtype=self.dtype_real, device="cuda")
        return torch.full(shp, float(value), dtype=self.dtype_real, device="cuda")

    def create_arrayK(self, value=None, shape=None):
        torch = _torch()
        shp = tuple(self.shapeK) if shape is None else tuple(shape)
        if value is None:
            return torch.empty(shp, dtype=self.dtype_complex, device="cuda")
        return torch.full(shp, complex(value), dtype=self.dtype_complex, device="cuda")

    def create_arrayX_random(self, min_val=None, max_val=None):
        torch = _torch()
        a = torch.rand(self.shapeX, dtype=self.dtype_real, device="cuda")
        if min_val is not None or max_val is not None:
            lo = 0.0 if min_val is None else min_val
            hi = 1.0 if max_val is None else max_val
            a = lo + (hi - lo) * a
        return a

    def create_arrayK_random(self):
        torch = _torch()
        a = torch.rand(self.shapeK, dtype=self.dtype_complex, device="cuda")
        return a

    # ------------------------------------------------------------ FFTs
    def fft_as_arg(self, field, field_fft):
        """Forward transform of field into field_fft (device tensors)."""
        f = self._to_dev(field, False)
        n = f.numel() // int(np.prod(self.shapeX))
        _lib.call("fsx_fft", self.plan.handle, _lib.dptr(f), _lib.dptr(field_fft),
                  n, _lib.stream_ptr())

    def ifft_as_arg(self, field_fft, field):
        f = self._to_dev(field_fft, True)
        n = f.numel() // int(np.prod(self.shapeK))
        _lib.call("fsx_ifft", self.plan.handle, _lib.dptr(f), _lib.dptr(field),
                  n, _lib.stream_ptr())

    ifft_as_arg_destroy = ifft_as_arg

    def fft(self, field):
        was_np = self._is_np(field)
        f = self._to_dev(field, False)
        lead = tuple(f.shape[: f.dim() - len(self.shapeX)])
        out = self.create_arrayK(shape=lead + tuple(self.shapeK))
        self.fft_as_arg(f, out)
        return out.cpu().numpy() if was_np else out

    def ifft(self, field_fft):
        was_np = self._is_np(field_fft)
        f = self._to_dev(field_fft, True)
        lead = tuple(f.shape[: f.dim() - len(self.shapeK)])
        out = self.create_arrayX(shape=lead + tuple(self.shapeX))
        self.ifft_as_arg(f, out)
        return out.cpu().numpy() if was_np else out

    # ------------------------------------------------------------ dealiasing
    def dealiasing(self, *args):
        """Zero the dealiased modes in place (operators2d.py:200-220,
        operators3d.py:336-342)."""
        if not self._has_to_dealiase:
            return
        torch = _torch()
        for thing in args:
            if isinstance(thing, SetOfVariables):
                thing = thing.data
            if isinstance(thing, np.ndarray):
                where = self.where_dealiased.astype(bool)
                if thing.ndim == len(self.shapeK):
                    thing[where] = 0.0
                else:
                    for a in thing:
                        a[where] = 0.0
            elif isinstance(thing, torch.Tensor):
                if thing.dtype != self.dtype_complex or not thing.is_contiguous():
                    raise ValueError("dealiasing needs a contiguous engine-dtype tensor")
                n = thing.numel() // int(np.prod(self.shapeK))
                _lib.call("fsx_dealias", self.plan.handle, _lib.dptr(thing), n, 0,
                          _lib.stream_ptr())

    dealiasing_setofvar = dealiasing
    dealiasing_variable = dealiasing

    # ------------------------------------------------------------ sums
    def _weights(self):
        """Parseval weights over the kx index (1 on kx=0 and the even-nx
        Nyquist column, 2 elsewhere)."""
        w = np.full(self.shapeK[-1], 2.0)
        w[0] = 1.0
        if self.nx % 2 == 0:
            w[-1] = 1.0
        return w

    def sum_wavenumbers(self, A):
        if self._is_np(A) and not hasattr(A, "device"):
            A = np.asarray(A)
            s = (A.real.astype(np.float64) * self._weights()).sum()
            return float(s)
        torch = _torch()
        w = torch.from_numpy(self._weights()).to(A.device)
        Ar = A.real if A.is_complex() else A
        return float((Ar.to(torch.float64) * w).sum().item())

    def mean_space(self, arr):
        """fluidsim/operators/base.py:74-81 (single process)."""
        if self._is_np(arr):
            return float(np.mean(arr))
        return float(arr.to(_torch().float64).mean().item())

    # ------------------------------------------------------------ K arrays
    def _dev(self, name):
        """Device copy of a host wavenumber array (cached)."""
        cache = self.__dict__.setdefault("_dev_cache", {})
        if name not in cache:
            arr = getattr(self, name)
            cache[name] = _torch().from_numpy(np.ascontiguousarray(arr)).to(
                device="cuda", dtype=self.dtype_real
            )
        return cache[name]
