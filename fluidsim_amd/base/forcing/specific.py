"""Forcing makers on a coarse spectral box.

Mirrors fluidsim/base/forcing/specific.py (SpecificForcingPseudoSpectralCoarse
135-328, InScriptForcingPseudoSpectralCoarse 349-381, NormalizedForcing
414-726, RandomSimplePseudoSpectral 729-765,
TimeCorrelatedRandomPseudoSpectral 768-868) and
fluidsim/base/forcing/anisotropic.py:24-236.

Division of work:
* host (numpy, at set-up and once per forcing period): the coarse grid, the
  forced region COND_NO_F, the seeded random coarse arrays forcing0 /
  forcing1 (numpy's legacy global generator, as the reference) and their
  projection on real fields; per step only the scalar time coefficient
  0.5 (cos((t - t_last) pi / T) + 1) is computed here;
* device: the per-step normalisation (sums over the forced modes of the
  state, root of the 2nd-degree equation; k_forcing_norm) runs inside the
  first stage of the step after the CFL update, and the Runge-Kutta stages
  add the coarse box to the tendencies (fsx_forcing_*).

The coarse size follows the reference's fallback for a missing ``pulp``
(specific.py:196-201): 2 x the smallest power of two >= nkmax_forcing.
fluidfft's create_arrayK_random is restated as independent uniform real
and imaginary parts on [min_val, 1).
"""

import ctypes
from math import radians
from pathlib import Path

import numpy as np

from ... import _lib

ROOTS = {"minabs": 0, "first": 1, "second": 2, "positive": 3}


def coarse_size(nk):
    """specific.py:196-201: 2^i >= 2 nk."""
    i = 0
    while 2 * nk > 2**i:
        i += 1
    return 2**i


def ensure_radians(angle):
    """fluidsim/util/util.py:722-732."""
    if isinstance(angle, str):
        if angle.endswith("°"):
            return radians(float(angle[:-1]))
        raise ValueError(
            "Angle should be a string with \nthe degree symbol or a float in radians"
        )
    return angle


class CoarseGrid:
    """Wavenumbers, Parseval weights and real projection of a small
    (coarse) spectral grid with the lengths of the simulation."""

    def __init__(self, sizes, lengths):
        self.sizes = list(sizes)  # (nx, ny[, nz])
        self.ndim = len(sizes)
        nx = sizes[0]
        ks = [2 * np.pi / L for L in lengths]
        kx = ks[0] * np.arange(nx // 2 + 1)

        def signed(n, dk):
            i = np.arange(n)
            return dk * np.where(i <= n // 2, i, i - n)

        if self.ndim == 2:
            ky = signed(sizes[1], ks[1])
            self.KX, self.KY = np.meshgrid(kx, ky)
            self.K2 = self.KX**2 + self.KY**2
            self.shapeK = (sizes[1], nx // 2 + 1)
            self.shapeX = (sizes[1], nx)
        else:
            ky = signed(sizes[1], ks[1])
            kz = signed(sizes[2], ks[2])
            self.Kz, self.Ky, self.Kx = np.meshgrid(kz, ky, kx, indexing="ij")
            self.K2 = self.Kx**2 + self.Ky**2 + self.Kz**2
            self.shapeK = (sizes[2], sizes[1], nx // 2 + 1)
            self.shapeX = (sizes[2], sizes[1], nx)
        self.K = np.sqrt(self.K2)

    def weights(self):
        """sum_wavenumbers weights: 1 on kx = 0 (and on the Nyquist column,
        even nx), 2 elsewhere."""
        w = np.full(self.shapeK, 2.0)
        w[..., 0] = 1.0
        if self.sizes[0] % 2 == 0:
            w[..., -1] = 1.0
        return w

    def project_fft_on_realX(self, f):
        """fluidsim/operators/operators2d.py:222-259 (2D);
        operators3d.py:1051-1052 (3D: fft(ifft(f)))."""
        if self.ndim == 3:
            x = np.fft.irfftn(f, s=self.shapeX, axes=(0, 1, 2), norm="forward")
            return np.fft.rfftn(x, axes=(0, 1, 2), norm="forward")
        f = f.copy()
        nky = self.shapeK[0]
        kyM = nky // 2
        kxM = self.shapeK[1] - 1
        for iy in (0, kyM):
            for ix in (0, kxM):
                f[iy, ix] = f[iy, ix].real
        for ikyp in range(1, kyM):
            ikyn = nky - ikyp
            for ix in (0, kxM):
                p, n = f[ikyp, ix], f[ikyn, ix]
                f[ikyp, ix] = (p + n.conjugate()) / 2
                f[ikyn, ix] = ((p + n.conjugate()) / 2).conjugate()
        return f


class SpecificForcing:
    tag = "specific"

    @classmethod
    def _complete_params_with_default(cls, params):
        params.forcing.available_types.append(cls.tag)

    def __init__(self, sim):
        self.sim = sim
        self.oper = sim.oper
        self.params = sim.params


class SpecificForcingPseudoSpectralCoarse(SpecificForcing):
    """specific.py:135-328: forcing defined on a coarse grid."""

    tag = "pseudo_spectral"
    _key_forced_default = "rot_fft"

    def __init__(self, sim):
        super().__init__(sim)
        params = sim.params
        pf = params.forcing
        if pf.nkmax_forcing < pf.nkmin_forcing:
            raise ValueError(
                f"params.forcing.nkmax_forcing = {pf.nkmax_forcing} < "
                f"params.forcing.nkmin_forcing = {pf.nkmin_forcing}"
            )
        self.kmax_forcing = self.oper.deltak * pf.nkmax_forcing
        self.kmin_forcing = self.oper.deltak * pf.nkmin_forcing
        self.forcing_rate = pf.forcing_rate
        self.key_forced = pf.key_forced if pf.key_forced is not None else (
            self._key_forced_default
        )
        sizes = self._coarse_sizes(coarse_size(int(round(pf.nkmax_forcing))))
        full = [self.oper.nx, self.oper.ny] + (
            [self.oper.nz] if self.oper.dim == 3 else []
        )
        if any(c > n for c, n in zip(sizes, full)):
            raise NotImplementedError(
                "The resolution is too small for the required forcing: "
                f"any(np.greater({sizes}, {full}))"
            )
        lengths = [self.oper.Lx, self.oper.Ly] + (
            [self.oper.Lz] if self.oper.dim == 3 else []
        )
        self.oper_coarse = CoarseGrid(sizes, lengths)
        self.shapeK_loc_coarse = self.oper_coarse.shapeK
        self.COND_NO_F = self._compute_cond_no_forcing()
        self.nb_forced_modes = int(self.COND_NO_F.size - self.COND_NO_F.sum())
        if not self.nb_forced_modes:
            raise ValueError("0 modes forced.")
        self.ind_forcing = np.logical_not(self.COND_NO_F).flatten().nonzero()[0]
        self._check_kept()
        self._setup_engine()

    # ------------------------------------------------------------ geometry
    def _coarse_sizes(self, size):
        return [size] * self.oper.dim

    def _compute_cond_no_forcing(self):
        """specific.py:249-265."""
        oc = self.oper_coarse
        C = np.logical_or(oc.K > self.kmax_forcing, oc.K < self.kmin_forcing)
        self._mask_nyquist(C)
        return C

    def _mask_nyquist(self, C):
        shp = self.oper_coarse.shapeK
        if self.oper.dim == 2:
            C[shp[0] // 2, :] = True
            C[:, shp[1] - 1] = True
        else:
            C[shp[0] // 2, :, :] = True
            C[:, shp[1] // 2, :] = True
            C[:, :, shp[2] - 1] = True

    def _full_indices(self):
        """Full-grid (iz, iy, ikx) of every coarse mode
        (put_coarse_array_in_array_fft, operators2d.py:525-600)."""
        shp = self.oper_coarse.shapeK
        nkxc = shp[-1]
        full_shape = self.oper.shapeK

        def mapi(nc, n):
            i = np.arange(nc)
            return np.where(i <= nc / 2, i, i - nc + n)

        if self.oper.dim == 2:
            iy = mapi(shp[0], full_shape[0])
            IY, IX = np.meshgrid(iy, np.arange(nkxc), indexing="ij")
            return (IY, IX)
        iz = mapi(shp[0], full_shape[0])
        iy = mapi(shp[1], full_shape[1])
        IZ, IY, IX = np.meshgrid(iz, iy, np.arange(nkxc), indexing="ij")
        return (IZ, IY, IX)

    def _check_kept(self):
        """Forced modes must survive the dealiasing of the tendencies."""
        wd = np.asarray(self.oper.where_dealiased, dtype=bool)
        idx = self._full_indices()
        forced = np.zeros(self.COND_NO_F.size, dtype=bool)
        forced[self._device_modes()] = True
        if np.any(wd[idx][forced.reshape(self.COND_NO_F.shape)]):
            raise ValueError("forced wavenumbers fall in the dealiased region")

    # ------------------------------------------------------------ engine
    def _key_kind(self):
        keys = list(self.sim.state.keys_state_spect)
        if self.key_forced in keys:
            return 0, keys.index(self.key_forced)
        if self.key_forced == "vt_fft" and self.oper.dim == 3:
            return 1, 0
        raise ValueError(f"cannot force the key {self.key_forced!r}")

    def _device_modes(self):
        """Forced modes plus, on the kx = 0 plane, their conjugate partners
        (-ky, -kz): the projection on real fields makes the random arrays
        Hermitian there, so a forcing region that is not symmetric (the
        anisotropic ones) has non-zero values outside ~COND_NO_F."""
        shp = self.oper_coarse.shapeK
        forced = ~self.COND_NO_F
        ext = forced.copy()
        if self.oper.dim == 2:
            ny = shp[0]
            iy = np.nonzero(forced[:, 0])[0]
            ext[(-iy) % ny, 0] = True
        else:
            nz, ny = shp[0], shp[1]
            iz, iy = np.nonzero(forced[:, :, 0])
            ext[(-iz) % nz, (-iy) % ny, 0] = True
        return np.nonzero(ext.ravel())[0]

    def _mode_weights(self):
        return self.oper_coarse.weights().ravel()[self._dev_modes]

    def _setup_engine(self):
        oc = self.oper_coarse
        shp = oc.shapeK
        nzc = shp[0] if self.oper.dim == 3 else 1
        nyc = shp[-2]
        nkxc = shp[-1]
        kind, var = self._key_kind()
        self._kind = kind
        self._dev_modes = self._device_modes()
        modes = np.ascontiguousarray(self._dev_modes, dtype=np.int32)
        w = np.ascontiguousarray(self._mode_weights(), dtype=np.float64)
        self._nvar = len(self.sim.state.keys_state_spect)
        _lib.call(
            "fsx_forcing_setup", self.oper.plan.handle, nyc, nzc, nkxc,
            self._nvar, len(modes), modes.ctypes.data_as(ctypes.c_void_p),
            w.ctypes.data_as(ctypes.c_void_p), kind, var,
        )
        _lib.call("fsx_forcing_enable", self.oper.plan.handle, 1)

    def _upload_pair(self, f0, f1):
        a0 = np.ascontiguousarray(f0.ravel()[self._dev_modes], dtype=np.complex128)
        a1 = np.ascontiguousarray(f1.ravel()[self._dev_modes], dtype=np.complex128)
        _lib.call(
            "fsx_forcing_set_pair", self.oper.plan.handle,
            a0.ctypes.data_as(ctypes.c_void_p), a1.ctypes.data_as(ctypes.c_void_p),
            _lib.stream_ptr(),
        )

    @property
    def forcing_fft(self):
        """Full-size forcing (put_forcingc_in_forcing, specific.py:279-290)."""
        from ...setofvariables import SetOfVariables

        out = SetOfVariables(like=self.sim.state.state_spect, info="forcing_fft")
        _lib.call(
            "fsx_forcing_add", self.oper.plan.handle, _lib.dptr(out.data),
            self._nvar, 1.0, _lib.stream_ptr(),
        )
        return out


class InScriptForcingPseudoSpectralCoarse(SpecificForcingPseudoSpectralCoarse):
    """specific.py:349-381: a coarse forcing given by the user script,
    constant until ``set_forcingc`` is called again (with the coarse
    arrays of the state keys, shape (nvar, *shapeK_loc_coarse))."""

    tag = "in_script_coarse"

    def __init__(self, sim):
        super().__init__(sim)
        self.is_initialized = False

    def _key_kind(self):
        # not normalised: the key only names what the script forces
        return 0, 0

    def set_forcingc(self, fc):
        fc = np.ascontiguousarray(fc, dtype=self.oper.np_complex)
        if fc.shape != (self._nvar,) + tuple(self.shapeK_loc_coarse):
            raise ValueError(f"coarse forcing must have shape "
                             f"{(self._nvar,) + tuple(self.shapeK_loc_coarse)}")
        # put_coarse_array_in_array_fft (operators2d.py:586-590)
        nyq = np.zeros(self.shapeK_loc_coarse, dtype=bool)
        self._mask_nyquist(nyq)
        if np.any(fc[:, nyq] != 0):
            raise ValueError("the coarse Nyquist modes of the forcing must be 0")
        _lib.call("fsx_forcing_set_coarse", self.oper.plan.handle,
                  fc.ctypes.data_as(ctypes.c_void_p), _lib.stream_ptr())
        self.is_initialized = True

    def compute(self):
        pass


class NormalizedForcing(SpecificForcingPseudoSpectralCoarse):
    """specific.py:414-726 (type "2nd_degree_eq")."""

    tag = "normalized"

    @classmethod
    def _complete_params_with_default(cls, params):
        super()._complete_params_with_default(params)
        if "normalized" not in params.forcing._tag_children:
            params.forcing._set_child(
                "normalized",
                {"type": "2nd_degree_eq", "which_root": "minabs",
                 "constant_rate_of": None},
            )

    def __init__(self, sim):
        pn = sim.params.forcing.normalized
        if pn.type != "2nd_degree_eq":
            raise NotImplementedError(
                f"params.forcing.normalized.type = {pn.type!r} is not available"
            )
        if pn.which_root not in ROOTS:
            raise ValueError(
                "Not sure how to choose which root to normalize forcing with."
            )
        self._constant_rate_of = pn.constant_rate_of
        super().__init__(sim)

    def _mode_weights(self):
        """Weights of the sums a, b: Parseval weights; for ns2d with
        constant_rate_of energy the velocity is forced through rot_fft, so
        both sums carry 1 / k^2 (ForcingNS2D.compute_coef_ab_normalize,
        fluidsim/solvers/ns2d/forcing.py:41-64)."""
        w = super()._mode_weights()
        if self._constant_rate_of is None:
            return w
        if self._constant_rate_of not in ("energy", "energyK") or (
            self.oper.dim != 2 or self.key_forced != "rot_fft"
        ):
            raise NotImplementedError(
                "constant_rate_of is available for ns2d forced through rot_fft"
            )
        K2 = self.oper_coarse.K2.ravel()[self._dev_modes]
        return w / K2


class RandomSimplePseudoSpectral(NormalizedForcing):
    """specific.py:729-765."""

    tag = "random"

    @classmethod
    def _complete_params_with_default(cls, params):
        super()._complete_params_with_default(params)
        if "random" not in params.forcing._tag_children:
            params.forcing._set_child("random", {"only_positive": False})

    def __init__(self, sim):
        self._min_val = None if sim.params.forcing.random.only_positive else -1.0
        super().__init__(sim)

    def create_arrayK_random(self):
        """Uniform real and imaginary parts on [min_val, 1)."""
        shape = self.oper_coarse.shapeK
        lo = 0.0 if self._min_val is None else self._min_val
        re = np.random.random(shape)
        im = np.random.random(shape)
        return (lo + (1.0 - lo) * re) + 1j * (lo + (1.0 - lo) * im)

    def compute_forcingc_raw(self):
        f = self.create_arrayK_random()
        f[self.COND_NO_F] = 0.0
        return self.oper_coarse.project_fft_on_realX(f)


class TimeCorrelatedRandomPseudoSpectral(RandomSimplePseudoSpectral):
    """specific.py:768-868."""

    tag = "tcrandom"

    @classmethod
    def _complete_params_with_default(cls, params):
        super()._complete_params_with_default(params)
        if "tcrandom" not in params.forcing._tag_children:
            params.forcing._set_child(
                "tcrandom", {"time_correlation": "based_on_forcing_rate"}
            )

    def __init__(self, sim):
        super().__init__(sim)
        output = sim.output
        self._forcing_state_file_path = (
            Path(output.path_run) / "_forcing_state.txt"
            if getattr(output, "path_run", None) else None
        )
        if self._forcing_state_file_path and self._forcing_state_file_path.exists():
            with open(self._forcing_state_file_path) as file:
                lines = file.readlines()
            t_last_change, seed0, seed1 = lines[-1].split()
            self.t_last_change = float(t_last_change)
            self._seed0, self._seed1 = int(seed0), int(seed1)
        else:
            self.t_last_change = self.sim.time_stepping.t
            self._seed0 = np.random.randint(0, 2**31)
            self._seed1 = np.random.randint(0, 2**31)
            self._save_state()
        np.random.seed(self._seed0)
        self.forcing0 = self.compute_forcingc_raw()
        np.random.seed(self._seed1)
        self.forcing1 = self.compute_forcingc_raw()
        self._upload_pair(self.forcing0, self.forcing1)
        pforcing = self.params.forcing
        try:
            time_correlation = pforcing[self.tag].time_correlation
        except (AttributeError, KeyError):
            time_correlation = pforcing.tcrandom.time_correlation
        if time_correlation == "based_on_forcing_rate":
            self.period_change_f0f1 = self.forcing_rate ** (-1.0 / 3)
        else:
            self.period_change_f0f1 = time_correlation

    def _save_state(self):
        if not self.params.output.HAS_TO_SAVE or self._forcing_state_file_path is None:
            return
        with open(self._forcing_state_file_path, "w") as file:
            file.write(
                "# do not modify by hand\n# t_last_change seed0 seed1\n"
                f"{self.t_last_change} {self._seed0} {self._seed1}\n"
            )

    def forcingc_raw_each_time(self):
        """specific.py:822-840: new random array when the period ends."""
        tsim = self.sim.time_stepping.t
        if tsim - self.t_last_change >= self.period_change_f0f1:
            self.t_last_change = tsim
            self._seed0 = self._seed1
            self.forcing0 = self.forcing1
            self._seed1 = np.random.randint(0, 2**31)
            np.random.seed(self._seed1)
            self.forcing1 = self.compute_forcingc_raw()
            self._save_state()
            self._upload_pair(self.forcing0, self.forcing1)

    def time_coef(self):
        """forcingc_from_f0f1 (specific.py:852-868):
        f = f1 - coef (f1 - f0), coef = 0.5 (cos((t - t_last) pi/T) + 1)."""
        omega = np.pi / self.period_change_f0f1
        return 0.5 * (np.cos((self.sim.time_stepping.t - self.t_last_change) * omega) + 1)

    def forcingc_from_f0f1(self):
        return self.forcing1 - self.time_coef() * (self.forcing1 - self.forcing0)

    def compute(self):
        """NormalizedForcing.compute (specific.py:460-500): the raw forcing
        of this step is set here; the normalisation on the state and the
        step's dt runs on the device before the first stage algebra."""
        self.forcingc_raw_each_time()
        _lib.call(
            "fsx_forcing_step", self.oper.plan.handle, float(self.time_coef()),
            float(self.forcing_rate), ROOTS[self.params.forcing.normalized.which_root],
        )
        self._pending = True


class TimeCorrelatedRandomPseudoSpectralAnisotropic(TimeCorrelatedRandomPseudoSpectral):
    """fluidsim/base/forcing/anisotropic.py:24-236."""

    tag = "tcrandom_anisotropic"
    _delta_angle_default = None

    @classmethod
    def _complete_params_with_default(cls, params):
        super(TimeCorrelatedRandomPseudoSpectral, cls)._complete_params_with_default(
            params
        )
        if "tcrandom_anisotropic" not in params.forcing._tag_children:
            params.forcing._set_child(
                "tcrandom_anisotropic",
                {"angle": "45°", "delta_angle": cls._delta_angle_default,
                 "kz_negative_enable": False},
            )

    def _coarse_sizes(self, size):
        """anisotropic.py:72-137 (2 fftw_grid_size(m) -> coarse_size(m))."""
        pa = self.params.forcing.tcrandom_anisotropic
        self.angle = angle = ensure_radians(pa.angle)
        delta_angle = pa.delta_angle
        if delta_angle is not None:
            delta_angle = ensure_radians(delta_angle)
        if delta_angle is None:
            self.khmax_forcing = np.sin(angle) * self.kmax_forcing
            self.kvmax_forcing = np.cos(angle) * self.kmax_forcing
        else:
            self.khmin_forcing = np.sin(angle - 0.5 * delta_angle) * self.kmin_forcing
            self.kvmin_forcing = np.cos(angle + 0.5 * delta_angle) * self.kmin_forcing
            self.khmax_forcing = np.sin(angle + 0.5 * delta_angle) * self.kmax_forcing
            self.kvmax_forcing = np.cos(angle - 0.5 * delta_angle) * self.kmax_forcing
        o = self.oper
        kymax = self.khmax_forcing if o.dim == 3 else self.kvmax_forcing
        nx = coarse_size(int(self.khmax_forcing / o.deltakx) + 1)
        ny = coarse_size(int(kymax / o.deltaky) + 1)
        if o.dim == 2:
            return [nx, ny]
        return [nx, ny, coarse_size(int(self.kvmax_forcing / o.deltakz) + 1)]

    def _compute_cond_no_forcing(self):
        """anisotropic.py:139-236."""
        oc = self.oper_coarse
        pa = self.params.forcing.tcrandom_anisotropic
        angle = self.angle
        delta_angle = pa.delta_angle
        if delta_angle is not None:
            delta_angle = ensure_radians(delta_angle)
        kz_neg = pa.kz_negative_enable
        if self.oper.dim == 2:
            Kh, Kv = oc.KX, oc.KY
        else:
            Kh, Kv = np.sqrt(oc.Kx**2 + oc.Ky**2), oc.Kz
        if delta_angle is None:
            self.khmin_forcing = np.sin(angle) * self.kmin_forcing
            self.kvmin_forcing = np.cos(angle) * self.kmin_forcing
            c_kh = np.logical_or(Kh > self.khmax_forcing, Kh < self.khmin_forcing)
            c_kv = np.logical_or(Kv > self.kvmax_forcing, Kv < self.kvmin_forcing)
            if kz_neg:
                c_kv = np.logical_and(
                    c_kv,
                    np.logical_or(Kv < -self.kvmax_forcing, Kv > -self.kvmin_forcing),
                )
            C = np.logical_or(c_kh, c_kv)
        else:
            K = oc.K
            Knz = K.copy()
            Knz[Knz == 0] = 1e-14
            theta = np.arccos(Kv / Knz)
            c_k = np.logical_or(K > self.kmax_forcing, K < self.kmin_forcing)
            c_t = np.logical_or(theta > angle + 0.5 * delta_angle,
                                theta < angle - 0.5 * delta_angle)
            if kz_neg:
                c_t = np.logical_and(
                    c_t,
                    np.logical_or(theta < np.pi - angle - 0.5 * delta_angle,
                                  theta > np.pi - angle + 0.5 * delta_angle),
                )
            C = np.logical_or(c_k, c_t)
        self._mask_nyquist(C)
        return C


class TimeCorrelatedRandomPseudoSpectralAnisotropic3D(
    TimeCorrelatedRandomPseudoSpectralAnisotropic
):
    """anisotropic.py:498-512: delta_angle defaults to 10 degrees."""

    _delta_angle_default = "10°"
