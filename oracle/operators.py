"""FP64 numpy restatement of the pseudo-spectral operators used on the hot path.

Test infrastructure only (see ``oracle/__init__.py``).

Conventions (sequential fluidfft layout, restated from the reference call
sites; fluidfft itself is not in the tree):

* 2D: X arrays are (ny, nx), K arrays (ny, nx//2+1), axes ("y", "x")
  (fluidsim/operators/operators2d.py:117).
* 3D: X arrays are (nz, ny, nx), K arrays (nz, ny, nx//2+1), axes
  ("z", "y", "x") (fluidsim/operators/operators3d.py:209); dimX_K == (0, 1, 2)
  (fluidsim/operators/test/test_operators3d.py:542).
* Wavenumber of index i along a y/z axis is i for i <= n/2 else i - n, so the
  Nyquist index maps to +n/2 (fluidsim/operators/operators3d.py:1196-1199).
* fft divides by the number of grid points, ifft does not, so that
  ifft(fft(x)) == x and Parseval reads sum_wavenumbers(|f|^2) == mean(f^2)
  (fluidsim/base/state.py:385-392).
"""

import numpy as np
import scipy.fft as sfft

WORKERS = 1


def set_workers(n):
    """Number of threads scipy.fft may use (1 = fluidsim's sequential mode)."""
    global WORKERS
    WORKERS = int(n)


def _k_signed(n, deltak):
    """Wavenumbers of a complex (y or z) axis, Nyquist mapped to +n/2."""
    idx = np.arange(n)
    k = np.where(idx <= n // 2, idx, idx - n)
    return deltak * k.astype(np.float64)


class _OperBase:
    def _finish_common(self, coef_dealiasing, truncation_shape):
        # K2_not0, K4, K8 (fluidsim/operators/operators3d.py:290-299; K8 as
        # K4**2, the same product fluidsim's compute_freq_diss multiplies by)
        self.K4 = self.K2**2
        self.K8 = self.K4**2
        self.K2_not0 = self.K2.copy()
        self.K2_not0.flat[0] = 1e-14
        self.K = np.sqrt(self.K2)
        self.coef_dealiasing = coef_dealiasing
        self._has_to_dealiase = bool(coef_dealiasing)
        self.truncation_shape = truncation_shape
        self.where_dealiased = self._compute_where_dealiased()

    def _compute_where_dealiased(self):
        """Mask of dealiased modes.

        cubic: |k_i| >= c * dk_i * n_i / 2 along any axis (fluidfft rule, as
        used by fluidsim/operators/operators2d.py:200-220); spherical and
        no_multiple_aliases follow fluidsim/operators/base.py:47-72 and
        operators2d.py:195-198 / operators3d.py:280-288.
        """
        c = self.coef_dealiasing
        shape = self.K2.shape
        if not c:
            return np.zeros(shape, dtype=np.uint8)
        axes = self._axes_k()
        cubic = np.zeros(shape, dtype=bool)
        for Ki, dk, n in axes:
            cubic |= np.abs(Ki) >= c * dk * n / 2
        if self.truncation_shape == "cubic":
            mask = cubic
        elif self.truncation_shape == "spherical":
            kmax = c * self.deltakx * self.nx / 2
            mask = self.K2 >= kmax**2
        elif self.truncation_shape == "no_multiple_aliases":
            al = [np.abs(Ki) >= 2 / 3 * dk * n / 2 for Ki, dk, n in axes]
            if len(al) == 2:
                mask = al[0] & al[1]
            else:
                mask = (al[0] & al[1]) | (al[1] & al[2]) | (al[2] & al[0])
            kmax = c * self.deltakx * self.nx / 2
            mask = mask | (self.K2 >= kmax**2)
        else:
            raise ValueError(
                'truncation_shape must be "cubic", "spherical" or '
                '"no_multiple_aliases"'
            )
        return np.array(mask, dtype=np.uint8)

    def dealiasing(self, *arrays):
        """Zero the dealiased modes in place (operators2d.py:200-220,
        operators3d.py:38-71)."""
        if not self._has_to_dealiase:
            return
        where = self.where_dealiased.astype(bool)
        for a in arrays:
            if a.ndim == self.K2.ndim:
                a[where] = 0.0
            else:
                for i in range(a.shape[0]):
                    a[i][where] = 0.0

    def sum_wavenumbers(self, A):
        """Sum over the half spectrum with weight 1 on kx=0 (and on the
        Nyquist column when nx is even), 2 elsewhere (Parseval weights)."""
        A = np.asarray(A)
        s = A[..., 0].sum(dtype=np.float64)
        if self.nx % 2 == 0:
            s += 2.0 * A[..., 1:-1].sum(dtype=np.float64)
            s += A[..., -1].sum(dtype=np.float64)
        else:
            s += 2.0 * A[..., 1:].sum(dtype=np.float64)
        return float(np.real(s))

    def mean_space(self, arr):
        """fluidsim/operators/base.py:74-81 (sequential branch)."""
        return float(np.mean(arr))

    def fft(self, x):
        """Forward real transform normalised by the number of points."""
        return sfft.rfftn(
            np.asarray(x, dtype=np.float64), norm="forward", workers=WORKERS
        )

    def ifft(self, x_fft):
        """Unnormalised inverse real transform (inverse of :meth:`fft`)."""
        return sfft.irfftn(
            np.asarray(x_fft, dtype=np.complex128),
            s=self.shapeX,
            norm="forward",
            workers=WORKERS,
        )

    def fft_as_arg(self, x, out):
        out[...] = self.fft(x)

    def ifft_as_arg(self, x_fft, out):
        out[...] = self.ifft(x_fft)

    def create_arrayK(self, value=0.0):
        return np.full(self.shapeK, value, dtype=np.complex128)

    def create_arrayX(self, value=0.0):
        return np.full(self.shapeX, value, dtype=np.float64)


class Oper2D(_OperBase):
    """Restatement of OperatorsPseudoSpectral2D
    (fluidsim/operators/operators2d.py:84-259) on one process."""

    def __init__(
        self,
        nx,
        ny,
        Lx=2 * np.pi,
        Ly=2 * np.pi,
        coef_dealiasing=2.0 / 3,
        truncation_shape="cubic",
        NO_SHEAR_MODES=False,
        NO_KY0=False,
    ):
        self.nx, self.ny = int(nx), int(ny)
        self.Lx, self.Ly = float(Lx), float(Ly)
        self.shapeX = (self.ny, self.nx)
        self.shapeK = (self.ny, self.nx // 2 + 1)
        self.deltax = self.Lx / self.nx
        self.deltay = self.Ly / self.ny
        self.deltakx = 2 * np.pi / self.Lx
        self.deltaky = 2 * np.pi / self.Ly
        kx = self.deltakx * np.arange(self.nx // 2 + 1, dtype=np.float64)
        ky = _k_signed(self.ny, self.deltaky)
        self.KX, self.KY = np.meshgrid(kx, ky)
        self.K2 = self.KX**2 + self.KY**2
        self._finish_common(coef_dealiasing, truncation_shape)
        # operators2d.py:171-193
        if NO_SHEAR_MODES:
            self.where_dealiased = np.array(
                (np.abs(self.KX) == 0.0) | self.where_dealiased.astype(bool),
                dtype=np.uint8,
            )
        if NO_KY0:
            self.where_dealiased = np.array(
                (np.abs(self.KY) == 0.0) | self.where_dealiased.astype(bool),
                dtype=np.uint8,
            )

    def _axes_k(self):
        return [(self.KX, self.deltakx, self.nx), (self.KY, self.deltaky, self.ny)]

    fft2 = _OperBase.fft
    ifft2 = _OperBase.ifft

    def vecfft_from_rotfft(self, rot_fft):
        """(ux, uy) = (i ky, -i kx) rot / k^2, zero mean mode."""
        inv = 1.0 / self.K2_not0
        ux_fft = 1j * self.KY * inv * rot_fft
        uy_fft = -1j * self.KX * inv * rot_fft
        return ux_fft, uy_fft

    def rotfft_from_vecfft(self, ux_fft, uy_fft):
        return 1j * (self.KX * uy_fft - self.KY * ux_fft)

    def divfft_from_vecfft(self, ux_fft, uy_fft):
        return 1j * (self.KX * ux_fft + self.KY * uy_fft)

    def gradfft_from_fft(self, f_fft):
        return 1j * self.KX * f_fft, 1j * self.KY * f_fft

    def projection_perp(self, ux_fft, uy_fft):
        """In-place projection on the plane perpendicular to k."""
        tmp = (self.KX * ux_fft + self.KY * uy_fft) / self.K2_not0
        ux_fft -= self.KX * tmp
        uy_fft -= self.KY * tmp
        return ux_fft, uy_fft


class Oper3D(_OperBase):
    """Restatement of OperatorsPseudoSpectral3D
    (fluidsim/operators/operators3d.py:116-342, 732-992) on one process."""

    def __init__(
        self,
        nx,
        ny,
        nz,
        Lx=2 * np.pi,
        Ly=2 * np.pi,
        Lz=2 * np.pi,
        coef_dealiasing=2.0 / 3,
        truncation_shape="cubic",
        NO_SHEAR_MODES=False,
    ):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.shapeX = (self.nz, self.ny, self.nx)
        self.shapeK = (self.nz, self.ny, self.nx // 2 + 1)
        self.deltax = self.Lx / self.nx
        self.deltay = self.Ly / self.ny
        self.deltaz = self.Lz / self.nz
        self.deltakx = 2 * np.pi / self.Lx
        self.deltaky = 2 * np.pi / self.Ly
        self.deltakz = 2 * np.pi / self.Lz
        kx = self.deltakx * np.arange(self.nx // 2 + 1, dtype=np.float64)
        ky = _k_signed(self.ny, self.deltaky)
        kz = _k_signed(self.nz, self.deltakz)
        self.Kz, self.Ky, self.Kx = np.meshgrid(kz, ky, kx, indexing="ij")
        self.K2 = self.Kx**2 + self.Ky**2 + self.Kz**2
        self._finish_common(coef_dealiasing, truncation_shape)
        if NO_SHEAR_MODES:
            # operators3d.py:268-276
            self.where_dealiased = np.array(
                (self.Kx**2 + self.Ky**2 == 0.0)
                | self.where_dealiased.astype(bool),
                dtype=np.uint8,
            )
        self.inv_K_square_nozero = 1.0 / self.K2_not0

    def _axes_k(self):
        return [
            (self.Kx, self.deltakx, self.nx),
            (self.Ky, self.deltaky, self.ny),
            (self.Kz, self.deltakz, self.nz),
        ]

    fft3d = _OperBase.fft
    ifft3d = _OperBase.ifft

    def ifft_as_arg_destroy(self, x_fft, out):
        out[...] = self.ifft(x_fft)

    def rotfft_from_vecfft(self, vx_fft, vy_fft, vz_fft):
        """omega_hat = i k x v_hat."""
        Kx, Ky, Kz = self.Kx, self.Ky, self.Kz
        return (
            1j * (Ky * vz_fft - Kz * vy_fft),
            1j * (Kz * vx_fft - Kx * vz_fft),
            1j * (Kx * vy_fft - Ky * vx_fft),
        )

    def project_perpk3d(self, vx_fft, vy_fft, vz_fft):
        """In-place Leray projection v -= k (k.v) / k^2."""
        tmp = (
            self.Kx * vx_fft + self.Ky * vy_fft + self.Kz * vz_fft
        ) * self.inv_K_square_nozero
        vx_fft -= self.Kx * tmp
        vy_fft -= self.Ky * tmp
        vz_fft -= self.Kz * tmp

    def project_kradial3d(self, vx_fft, vy_fft, vz_fft):
        """operators3d.py:732-780."""
        K2nz = self.K2.copy()
        K2nz[K2nz == 0] = 1e-14
        tmp = (self.Kx * vx_fft + self.Ky * vy_fft + self.Kz * vz_fft) / K2nz
        vx_fft[...] = self.Kx * tmp
        vy_fft[...] = self.Ky * tmp
        vz_fft[...] = self.Kz * tmp

    def _angles(self):
        Kh2 = self.Kx**2 + self.Ky**2
        K2nz = Kh2 + self.Kz**2
        Kh2nz = Kh2.copy()
        Kh2nz[Kh2nz == 0] = 1e-14
        K2nz[K2nz == 0] = 1e-14
        cos_theta = self.Kz * np.sqrt(1.0 / K2nz)
        sin_theta = np.sqrt(Kh2 / K2nz)
        cos_phi = self.Kx * np.sqrt(1.0 / Kh2nz)
        sin_phi = self.Ky * np.sqrt(1.0 / Kh2nz)
        return cos_theta, sin_theta, cos_phi, sin_phi

    def project_poloidal(self, vx_fft, vy_fft, vz_fft):
        """operators3d.py:787-843."""
        ct, st, cp, sp = self._angles()
        tmp = ct * cp * vx_fft + ct * sp * vy_fft - st * vz_fft
        vx_fft[...] = ct * cp * tmp
        vy_fft[...] = ct * sp * tmp
        vz_fft[...] = -st * tmp

    def project_toroidal(self, vx_fft, vy_fft, vz_fft):
        """operators3d.py:893-937."""
        _, _, cp, sp = self._angles()
        tmp = -sp * vx_fft + cp * vy_fft
        vx_fft[...] = -sp * tmp
        vy_fft[...] = cp * tmp
        vz_fft[...] = 0.0

    def urudfft_from_vxvyfft(self, vx_fft, vy_fft):
        """Toroidal / poloidal split of the horizontal velocity
        (operators3d.py:713-730): ud = k_h (k_h . v_h) / k_h^2 with
        k_h^2 = 0 replaced by 1e-14, ur = v_h - ud."""
        kh2 = self.Kx**2 + self.Ky**2
        kh2[kh2 == 0] = 1e-14
        inv = 1 / kh2
        kdotu = self.Kx * vx_fft + self.Ky * vy_fft
        udx = kdotu * self.Kx * inv
        udy = kdotu * self.Ky * inv
        return vx_fft - udx, vy_fft - udy, udx, udy

    def vtfft_from_vecfft(self, vx_fft, vy_fft, vz_fft):
        """operators3d.py:939-955."""
        _, _, cp, sp = self._angles()
        return -1j * sp * vx_fft + 1j * cp * vy_fft

    def vecfft_from_vtfft(self, vt_fft):
        """operators3d.py:957-972."""
        _, _, cp, sp = self._angles()
        return 1j * sp * vt_fft, -1j * cp * vt_fft, np.zeros_like(vt_fft)

    @staticmethod
    def vector_product(ax, ay, az, bx, by, bz):
        """a x b (fluidfft vector_product, used as v x omega in
        fluidsim/solvers/ns3d/solver.py:226)."""
        return (ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx)

    def div_vb_fft_from_vb(self, vx, vy, vz, b):
        """i (kx fft(vx b) + ky fft(vy b) + kz fft(vz b))."""
        return 1j * (
            self.Kx * self.fft(vx * b)
            + self.Ky * self.fft(vy * b)
            + self.Kz * self.fft(vz * b)
        )
