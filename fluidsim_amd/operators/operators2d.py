"""2D pseudo-spectral operators on the MI355X engine.

Drop-in for fluidsim.operators.operators2d.OperatorsPseudoSpectral2D
(fluidsim/operators/operators2d.py:84-259): same parameters
(params.oper.nx, ny, Lx, Ly, coef_dealiasing, truncation_shape,
NO_SHEAR_MODES, NO_KY0, type_fft) plus params.oper.precision ("float64"
default, or "float32").  Layout: X (ny, nx), K (ny, nx//2+1).
"""

from functools import cached_property

import numpy as np

from ..params import Parameters
from .base import OperatorsBase, k_signed, _torch
from .. import _lib


class OperatorsPseudoSpectral2D(OperatorsBase):
    dim = 2

    @classmethod
    def _create_default_params(cls):
        params = Parameters(tag="params", attribs={"ONLY_COARSE_OPER": False})
        cls._complete_params_with_default(params)
        return params

    @staticmethod
    def _complete_params_with_default(params):
        """operators2d.py:99-113 (+ precision)."""
        attribs = {
            "type_fft": "default",
            "coef_dealiasing": 2.0 / 3,
            "nx": 48,
            "ny": 48,
            "Lx": 8,
            "Ly": 8,
            "truncation_shape": "cubic",
            "NO_SHEAR_MODES": False,
            "NO_KY0": False,
            "precision": "float64",
        }
        params._set_child("oper", attribs=attribs)

    def __init__(self, params):
        self.params = params
        self.axes = ("y", "x")
        nx, ny = int(params.oper.nx), int(params.oper.ny)
        if params.oper.nx != nx:
            raise ValueError(f"params.oper.nx != int(params.oper.nx); ({params.oper.nx})")
        if params.oper.ny != ny:
            raise ValueError(f"params.oper.ny != int(params.oper.ny); ({params.oper.ny})")
        params.oper.nx, params.oper.ny = nx, ny
        if getattr(params, "ONLY_COARSE_OPER", False):
            nx = ny = 4
        self._init_precision(params)
        self.nx, self.ny = nx, ny
        self.Lx = self.lx = float(params.oper.Lx)
        self.Ly = self.ly = float(params.oper.Ly)
        self.shapeX = self.shapeX_seq = self.shapeX_loc = (ny, nx)
        self.shapeK = self.shapeK_seq = self.shapeK_loc = (ny, nx // 2 + 1)
        self.deltax = self.Lx / nx
        self.deltay = self.Ly / ny
        self.deltakx = 2 * np.pi / self.Lx
        self.deltaky = 2 * np.pi / self.Ly
        self.is_sequential = True
        self.is_transposed = False
        self.coef_dealiasing = params.oper.coef_dealiasing
        self._has_to_dealiase = bool(self.coef_dealiasing)
        self.where_dealiased = self._compute_where_dealiased(
            params.oper.truncation_shape
        )
        # operators2d.py:171-193
        if getattr(params.oper, "NO_SHEAR_MODES", False):
            self.where_dealiased = np.array(
                (np.abs(self.KX) == 0.0) | self.where_dealiased.astype(bool),
                dtype=np.uint8,
            )
            self._has_to_dealiase = True
        if getattr(params.oper, "NO_KY0", False):
            self.where_dealiased = np.array(
                (np.abs(self.KY) == 0.0) | self.where_dealiased.astype(bool),
                dtype=np.uint8,
            )
            self._has_to_dealiase = True
        self._init_plan(nf_max=4)

    def _axes_k_1d(self):
        kx = self.deltakx * np.arange(self.nx // 2 + 1, dtype=np.float64)
        ky = k_signed(self.ny, self.deltaky)
        # axis order of the K array: (ky, kx)
        return [(ky, self.deltaky, self.ny), (kx, self.deltakx, self.nx)]

    def _K2_host(self):
        return self.K2

    @property
    def deltak(self):
        """Largest of the deltak_i (fluidsim/base/forcing/specific.py
        :178-182 uses it for the forcing wavenumbers)."""
        return max(self.deltakx, self.deltaky)

    @cached_property
    def kx(self):
        return self.deltakx * np.arange(self.nx // 2 + 1, dtype=np.float64)

    @cached_property
    def ky(self):
        return k_signed(self.ny, self.deltaky)

    @cached_property
    def KX(self):
        return np.broadcast_to(self.kx, self.shapeK).copy()

    @cached_property
    def KY(self):
        return np.broadcast_to(self.ky[:, None], self.shapeK).copy()

    @cached_property
    def K2(self):
        return self.KX**2 + self.KY**2

    @cached_property
    def K4(self):
        return self.K2**2

    @cached_property
    def K8(self):
        return self.K4**2

    @cached_property
    def K(self):
        return np.sqrt(self.K2)

    @cached_property
    def K2_not0(self):
        a = self.K2.copy()
        a[0, 0] = 1e-14
        return a

    @cached_property
    def x(self):
        return self.deltax * np.arange(self.nx)

    @cached_property
    def y(self):
        return self.deltay * np.arange(self.ny)

    @property
    def x_seq(self):
        return self.x

    x_loc = x_seq

    @property
    def y_seq(self):
        return self.y

    y_loc = y_seq

    def produce_str_describing_oper(self):
        def fmt(L):
            return repr(int(L / np.pi)) + "pi" if (L / np.pi).is_integer() else f"{L:.3f}".rstrip("0")
        return f"{self.nx}x{self.ny}_S{fmt(self.Lx)}x{fmt(self.Ly)}"

    # --- fluidfft 2D operator arithmetic (torch on device) --------------
    fft2 = OperatorsBase.fft
    ifft2 = OperatorsBase.ifft

    def _k(self):
        return self._dev("KX"), self._dev("KY")

    def _inv_k2(self):
        cache = self.__dict__.setdefault("_dev_cache", {})
        if "inv_K2_not0" not in cache:
            cache["inv_K2_not0"] = 1.0 / self._dev("K2_not0")
        return cache["inv_K2_not0"]

    def _wrap(self, fn, *arrays):
        was_np = any(self._is_np(a) for a in arrays)
        devs = [self._to_dev(a, True) for a in arrays]
        out = fn(*devs)
        if was_np:
            if isinstance(out, tuple):
                return tuple(o.cpu().numpy() for o in out)
            return out.cpu().numpy()
        return out

    def vecfft_from_rotfft(self, rot_fft):
        """(ux, uy) = (i ky, -i kx) rot / k^2."""
        def f(r):
            KX, KY = self._k()
            inv = self._inv_k2()
            return 1j * KY * inv * r, -1j * KX * inv * r
        return self._wrap(f, rot_fft)

    def rotfft_from_vecfft(self, ux_fft, uy_fft):
        def f(ux, uy):
            KX, KY = self._k()
            return 1j * (KX * uy - KY * ux)
        return self._wrap(f, ux_fft, uy_fft)

    def divfft_from_vecfft(self, ux_fft, uy_fft):
        def f(ux, uy):
            KX, KY = self._k()
            return 1j * (KX * ux + KY * uy)
        return self._wrap(f, ux_fft, uy_fft)

    def gradfft_from_fft(self, f_fft):
        def f(a):
            KX, KY = self._k()
            return 1j * KX * a, 1j * KY * a
        return self._wrap(f, f_fft)

    def laplacian_fft(self, a_fft, order=2, negative=False):
        def f(a):
            r = a * self._dev("K2") ** (order // 2)
            return -r if not negative else r
        return self._wrap(f, a_fft)

    def invlaplacian_fft(self, a_fft, order=2, negative=False):
        def f(a):
            r = a / self._dev("K2_not0") ** (order // 2)
            r[0, 0] = 0.0
            return -r if not negative else r
        return self._wrap(f, a_fft)

    def projection_perp(self, ux_fft, uy_fft):
        """In-place projection on the plane perpendicular to k."""
        torch = _torch()
        if isinstance(ux_fft, torch.Tensor):
            KX, KY = self._k()
            tmp = (KX * ux_fft + KY * uy_fft) * self._inv_k2()
            ux_fft -= KX * tmp
            uy_fft -= KY * tmp
            return ux_fft, uy_fft
        tmp = (self.KX * ux_fft + self.KY * uy_fft) / self.K2_not0
        ux_fft -= self.KX * tmp
        uy_fft -= self.KY * tmp
        return ux_fft, uy_fft

    # --- fused engine entry -------------------------------------------------
    def ns2d_tendency(self, rot_fft, out_fft, beta=0.0, want_maxima=False):
        """Fused N = dealias(fft(-u.grad(rot) - beta v)) on the device."""
        _lib.call("fsx_ns2d_tendency", self.plan.handle, _lib.dptr(rot_fft),
                  _lib.dptr(out_fft), float(beta), int(want_maxima),
                  _lib.stream_ptr())
