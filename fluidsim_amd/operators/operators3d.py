"""3D pseudo-spectral operators on the MI355X engine.

Drop-in for fluidsim.operators.operators3d.OperatorsPseudoSpectral3D
(fluidsim/operators/operators3d.py:116-342, 732-992): parameters
params.oper.{nx, ny, nz, Lx, Ly, Lz, coef_dealiasing, truncation_shape,
NO_SHEAR_MODES, type_fft, type_fft2d} plus params.oper.precision.
Layout: X (nz, ny, nx), K (nz, ny, nx//2+1), dimX_K == (0, 1, 2).
Host wavenumber arrays (Kx, K2, ...) are built lazily: the engine kernels
compute wavenumbers from indices and never read them.
"""

from functools import cached_property

import numpy as np

from ..params import Parameters
from .base import OperatorsBase, k_signed, _torch
from .. import _lib

PROJ = {None: 0, "perpk": 1, "toroidal": 2, "vortical": 2, "poloidal": 3}


class OperatorsPseudoSpectral3D(OperatorsBase):
    dim = 3

    @classmethod
    def _create_default_params(cls):
        params = Parameters(tag="params", attribs={"ONLY_COARSE_OPER": False})
        cls._complete_params_with_default(params)
        return params

    @staticmethod
    def _complete_params_with_default(params):
        """operators3d.py:134-178 (+ precision)."""
        attribs = {
            "type_fft": "default",
            "type_fft2d": "sequential",
            "coef_dealiasing": 2.0 / 3,
            "nx": 48,
            "ny": 48,
            "nz": 48,
            "Lx": 2 * np.pi,
            "Ly": 2 * np.pi,
            "Lz": 2 * np.pi,
            "truncation_shape": "cubic",
            "NO_SHEAR_MODES": False,
            "precision": "float64",
        }
        params._set_child("oper", attribs=attribs)

    def __init__(self, params):
        self.params = params
        self.axes = ("z", "y", "x")
        for k in ("nx", "ny", "nz"):
            v = int(params.oper[k])
            if params.oper[k] != v:
                raise ValueError(f"params.oper.{k} != int(params.oper.{k})")
            params.oper[k] = v
        nx, ny, nz = params.oper.nx, params.oper.ny, params.oper.nz
        if getattr(params, "ONLY_COARSE_OPER", False):
            nx = ny = nz = 4
        self._init_precision(params)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx = float(params.oper.Lx)
        self.Ly = float(params.oper.Ly)
        self.Lz = float(params.oper.Lz)
        self.shapeX = self.shapeX_seq = self.shapeX_loc = (nz, ny, nx)
        self.shapeK = self.shapeK_seq = self.shapeK_loc = (nz, ny, nx // 2 + 1)
        self.dimX_K = (0, 1, 2)
        self.seq_indices_first_K = (0, 0, 0)
        self.deltax = self.Lx / nx
        self.deltay = self.Ly / ny
        self.deltaz = self.Lz / nz
        self.deltakx = 2 * np.pi / self.Lx
        self.deltaky = 2 * np.pi / self.Ly
        self.deltakz = 2 * np.pi / self.Lz
        self.is_sequential = True
        self.coef_dealiasing = params.oper.coef_dealiasing
        self._has_to_dealiase = bool(self.coef_dealiasing)
        self.where_dealiased = self._compute_where_dealiased(
            params.oper.truncation_shape
        )
        if getattr(params.oper, "NO_SHEAR_MODES", False):
            kx0 = (self.kx == 0)[None, None, :]
            ky0 = (self.ky == 0)[None, :, None]
            self.where_dealiased = np.array(
                (kx0 & ky0) | self.where_dealiased.astype(bool), dtype=np.uint8
            )
            self._has_to_dealiase = True
        self._init_plan(nf_max=7)

    def _axes_k_1d(self):
        return [
            (self.kz, self.deltakz, self.nz),
            (self.ky, self.deltaky, self.ny),
            (self.kx, self.deltakx, self.nx),
        ]

    def _K2_host(self):
        return self.K2

    @property
    def deltak(self):
        """Largest of the deltak_i (fluidsim/base/forcing/specific.py
        :178-182 uses it for the forcing wavenumbers)."""
        return max(self.deltakx, self.deltaky, self.deltakz)

    @cached_property
    def kx(self):
        return self.deltakx * np.arange(self.nx // 2 + 1, dtype=np.float64)

    @cached_property
    def ky(self):
        return k_signed(self.ny, self.deltaky)

    @cached_property
    def kz(self):
        return k_signed(self.nz, self.deltakz)

    @cached_property
    def Kx(self):
        return np.broadcast_to(self.kx[None, None, :], self.shapeK).copy()

    @cached_property
    def Ky(self):
        return np.broadcast_to(self.ky[None, :, None], self.shapeK).copy()

    @cached_property
    def Kz(self):
        return np.broadcast_to(self.kz[:, None, None], self.shapeK).copy()

    @cached_property
    def K2(self):
        return (self.kx[None, None, :] ** 2 + self.ky[None, :, None] ** 2
                + self.kz[:, None, None] ** 2)

    @cached_property
    def K4(self):
        return self.K2**2

    @cached_property
    def K8(self):
        return self.K4**2

    @cached_property
    def K2_not0(self):
        a = self.K2.copy()
        a[0, 0, 0] = 1e-14
        return a

    @cached_property
    def inv_K_square_nozero(self):
        return 1.0 / self.K2_not0

    @cached_property
    def x_seq(self):
        return self.deltax * np.arange(self.nx)

    @cached_property
    def y_seq(self):
        return self.deltay * np.arange(self.ny)

    @cached_property
    def z_seq(self):
        return self.deltaz * np.arange(self.nz)

    def produce_str_describing_oper(self):
        def fmt(L):
            return repr(int(L / np.pi)) + "pi" if (L / np.pi).is_integer() else f"{L:.3f}".rstrip("0")
        return (f"{self.nx}x{self.ny}x{self.nz}_V{fmt(self.Lx)}x"
                f"{fmt(self.Ly)}x{fmt(self.Lz)}")

    fft3d = OperatorsBase.fft
    ifft3d = OperatorsBase.ifft

    # --- fluidfft 3D operator arithmetic -------------------------------
    def _kdev(self):
        torch = _torch()
        cache = self.__dict__.setdefault("_dev_cache", {})
        if "k1d" not in cache:
            def d(a):
                return torch.from_numpy(a).to(device="cuda", dtype=self.dtype_real)
            cache["k1d"] = (
                d(self.kx)[None, None, :], d(self.ky)[None, :, None],
                d(self.kz)[:, None, None],
            )
        return cache["k1d"]

    def _wrap(self, fn, *arrays):
        was_np = any(self._is_np(a) for a in arrays)
        devs = [self._to_dev(a, True) for a in arrays]
        out = fn(*devs)
        if was_np:
            if isinstance(out, tuple):
                return tuple(o.cpu().numpy() for o in out)
            return out.cpu().numpy()
        return out

    def rotfft_from_vecfft(self, vx_fft, vy_fft, vz_fft):
        """omega = i k x v."""
        def f(vx, vy, vz):
            Kx, Ky, Kz = self._kdev()
            return (1j * (Ky * vz - Kz * vy), 1j * (Kz * vx - Kx * vz),
                    1j * (Kx * vy - Ky * vx))
        return self._wrap(f, vx_fft, vy_fft, vz_fft)

    def rotfft_from_vecfft_outin(self, vx_fft, vy_fft, vz_fft, rx, ry, rz):
        ox, oy, oz = self.rotfft_from_vecfft(vx_fft, vy_fft, vz_fft)
        rx[...] = ox
        ry[...] = oy
        rz[...] = oz

    def divfft_from_vecfft(self, vx_fft, vy_fft, vz_fft):
        def f(vx, vy, vz):
            Kx, Ky, Kz = self._kdev()
            return 1j * (Kx * vx + Ky * vy + Kz * vz)
        return self._wrap(f, vx_fft, vy_fft, vz_fft)

    def _project(self, vx_fft, vy_fft, vz_fft, kind):
        torch = _torch()
        if isinstance(vx_fft, np.ndarray):
            arrs = [self._to_dev(a, True) for a in (vx_fft, vy_fft, vz_fft)]
            _lib.call("fsx_project", self.plan.handle, *(_lib.dptr(a) for a in arrs),
                      kind, _lib.stream_ptr())
            for dst, src in zip((vx_fft, vy_fft, vz_fft), arrs):
                dst[...] = src.cpu().numpy()
            return
        for a in (vx_fft, vy_fft, vz_fft):
            if not (isinstance(a, torch.Tensor) and a.is_contiguous()
                    and a.dtype == self.dtype_complex):
                raise ValueError("projection needs contiguous engine-dtype tensors")
        _lib.call("fsx_project", self.plan.handle, _lib.dptr(vx_fft),
                  _lib.dptr(vy_fft), _lib.dptr(vz_fft), kind, _lib.stream_ptr())

    def project_perpk3d(self, vx_fft, vy_fft, vz_fft):
        """In-place Leray projection v -= k (k.v)/k^2."""
        self._project(vx_fft, vy_fft, vz_fft, 1)

    def project_toroidal(self, vx_fft, vy_fft, vz_fft):
        """operators3d.py:893-937."""
        self._project(vx_fft, vy_fft, vz_fft, 2)

    def project_poloidal(self, vx_fft, vy_fft, vz_fft):
        """operators3d.py:787-843."""
        self._project(vx_fft, vy_fft, vz_fft, 3)

    def _angles(self):
        Kx, Ky, Kz = self._kdev()
        kh2 = Kx**2 + Ky**2
        kh2n = kh2.clone()
        kh2n[kh2n == 0] = 1e-14
        return Kx / kh2n.sqrt(), Ky / kh2n.sqrt()

    def vtfft_from_vecfft(self, vx_fft, vy_fft, vz_fft):
        """operators3d.py:939-955."""
        def f(vx, vy, vz):
            cp, sp = self._angles()
            return -1j * sp * vx + 1j * cp * vy
        return self._wrap(f, vx_fft, vy_fft, vz_fft)

    def vecfft_from_vtfft(self, vt_fft):
        """operators3d.py:957-972."""
        torch = _torch()

        def f(vt):
            cp, sp = self._angles()
            return 1j * sp * vt, -1j * cp * vt, torch.zeros_like(vt)
        return self._wrap(f, vt_fft)

    @staticmethod
    def vector_product(ax, ay, az, bx, by, bz):
        return (ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx)

    def div_vb_fft_from_vb(self, vx, vy, vz, b):
        def f(a, bb, c, d):
            Kx, Ky, Kz = self._kdev()
            return 1j * (Kx * self.fft(a * d) + Ky * self.fft(bb * d)
                         + Kz * self.fft(c * d))
        was_np = any(self._is_np(a) for a in (vx, vy, vz, b))
        devs = [self._to_dev(a, False) for a in (vx, vy, vz, b)]
        out = f(*devs)
        return out.cpu().numpy() if was_np else out

    # --- fused engine entry -------------------------------------------------
    def ns3d_tendency(self, v_fft, out_fft, f=None, want_maxima=False):
        """Fused fft(v x omega), dealiased, not projected (3 fields)."""
        _lib.call("fsx_ns3d_tendency", self.plan.handle, _lib.dptr(v_fft),
                  _lib.dptr(out_fft), int(f is not None),
                  float(f) if f is not None else 0.0, int(want_maxima),
                  _lib.stream_ptr())
