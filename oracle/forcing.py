"""FP64 restatement of fluidsim's time-correlated random forcing ("tcrandom"
and "tcrandom_anisotropic") on a coarse spectral grid.

Test infrastructure only (see ``oracle/__init__.py``).

Reference: fluidsim/base/forcing/specific.py
  SpecificForcingPseudoSpectralCoarse 135-328 (coarse grid, COND_NO_F,
  put_forcingc_in_forcing), NormalizedForcing 414-726 (2nd-degree-equation
  normalisation, root choice), RandomSimplePseudoSpectral 729-765,
  TimeCorrelatedRandomPseudoSpectral 768-868;
fluidsim/base/forcing/anisotropic.py 24-236 (forcing region of the
anisotropic variant);
fluidsim/operators/operators2d.py 222-259, 264-281, 525-600 (real
projection, coarse <-> full copies).

The coarse size uses the reference's fallback rule for a missing ``pulp``
(specific.py:196-201): the smallest power of two >= 2 nkmax_forcing.

fluidfft's ``create_arrayK_random`` is absent; the random coarse array is
restated as independent uniform real and imaginary parts on
[min_val, 1) drawn with numpy's legacy global generator (real parts first).
"""

import numpy as np

from .operators import Oper2D, Oper3D


def coarse_size(nk):
    """specific.py:196-201 (power of two >= 2 nk)."""
    i = 0
    while 2 * nk > 2**i:
        i += 1
    return 2**i


def roots_2nd_degree(a, b, c, which_root="minabs"):
    """coef_normalization_from_abc (specific.py:679-726): roots of
    a R^2 + b R + c = 0, "first" = (-b + sqrt(D)) / 2a, "second" =
    (-b - sqrt(D)) / 2a, computed without cancellation (q = -(b + sgn(b)
    sqrt(D)) / 2, roots q / a and c / q).  A negative discriminant gives the
    complex pair with the positive imaginary part first (numpy's order)."""
    if a == 0.0:
        return 0.0  # np.roots gives one root: the reference returns 0
    delta = b * b - 4 * a * c
    if delta >= 0:
        sq = np.sqrt(delta)
        if b >= 0:
            q = -0.5 * (b + sq)
            second, first = q / a, (c / q if q != 0 else 0.0)
        else:
            q = -0.5 * (b - sq)
            first, second = q / a, c / q
    else:
        sq = np.sqrt(-delta)
        first = complex(-b / (2 * a), sq / (2 * a))
        second = first.conjugate()
    r1, r2 = first, second
    if which_root == "minabs":
        return r2 if abs(r2) < abs(r1) else r1
    if which_root == "first":
        return r1
    if which_root == "second":
        return r2
    if which_root == "positive":
        return r2 if np.real(r2) > 0.0 else r1
    raise ValueError(which_root)


class TCRandomForcing:
    """Time-correlated random normalised forcing of one key.

    ``compute(state_spect, t, dt)`` returns the full-grid forcing (nvar,
    *shapeK) that fluidsim adds to the tendencies at every stage of the
    step (computed once per step after the CFL update, base.py:208-214).
    """

    def __init__(
        self,
        oper,
        keys_state_spect,
        nkmin_forcing=4,
        nkmax_forcing=5,
        forcing_rate=1.0,
        key_forced=None,
        which_root="minabs",
        time_correlation="based_on_forcing_rate",
        only_positive=False,
        t0=0.0,
        anisotropic=None,
    ):
        self.oper = oper
        self.keys = tuple(keys_state_spect)
        self.ndim = oper.K2.ndim
        self.key_forced = key_forced or "rot_fft"
        self.forcing_rate = float(forcing_rate)
        self.which_root = which_root
        self.min_val = None if only_positive else -1.0
        self.anisotropic = anisotropic
        if nkmax_forcing < nkmin_forcing:
            raise ValueError("nkmax_forcing < nkmin_forcing")
        deltak = max(
            oper.deltakx, oper.deltaky, getattr(oper, "deltakz", 0.0)
        )
        self.kmax_forcing = deltak * nkmax_forcing
        self.kmin_forcing = deltak * nkmin_forcing
        size = coarse_size(nkmax_forcing)
        sizes = [size] * self.ndim
        if anisotropic is not None:
            sizes = self._anisotropic_sizes(anisotropic)
        if any(s > n for s, n in zip(sizes, self._full_sizes())):
            raise NotImplementedError("resolution too small for the forcing")
        if self.ndim == 2:
            self.oper_c = Oper2D(sizes[0], sizes[1], oper.Lx, oper.Ly,
                                 coef_dealiasing=1.0)
        else:
            self.oper_c = Oper3D(sizes[0], sizes[1], sizes[2], oper.Lx,
                                 oper.Ly, oper.Lz, coef_dealiasing=1.0)
        self.COND_NO_F = self._compute_cond_no_forcing()
        if not np.sum(~self.COND_NO_F):
            raise ValueError("0 modes forced.")
        self.t_last_change = t0
        self._seed0 = np.random.randint(0, 2**31)
        self._seed1 = np.random.randint(0, 2**31)
        np.random.seed(self._seed0)
        self.forcing0 = self.compute_forcingc_raw()
        np.random.seed(self._seed1)
        self.forcing1 = self.compute_forcingc_raw()
        if time_correlation == "based_on_forcing_rate":
            self.period_change_f0f1 = self.forcing_rate ** (-1.0 / 3)
        else:
            self.period_change_f0f1 = float(time_correlation)

    # ------------------------------------------------------------ geometry
    def _full_sizes(self):
        o = self.oper
        return [o.nx, o.ny] if self.ndim == 2 else [o.nx, o.ny, o.nz]

    def _anisotropic_sizes(self, an):
        """anisotropic.py:72-137 with the power-of-two size rule."""
        angle = an["angle"]
        delta_angle = an.get("delta_angle")
        if delta_angle is None:
            self.khmax_forcing = np.sin(angle) * self.kmax_forcing
            self.kvmax_forcing = np.cos(angle) * self.kmax_forcing
        else:
            self.khmin_forcing = np.sin(angle - 0.5 * delta_angle) * self.kmin_forcing
            self.kvmin_forcing = np.cos(angle + 0.5 * delta_angle) * self.kmin_forcing
            self.khmax_forcing = np.sin(angle + 0.5 * delta_angle) * self.kmax_forcing
            self.kvmax_forcing = np.cos(angle - 0.5 * delta_angle) * self.kmax_forcing
        o = self.oper
        kymax = self.khmax_forcing if self.ndim == 3 else self.kvmax_forcing
        # 2 fftw_grid_size(m) -> 2 x (power of two >= m) = coarse_size(m)
        nx = coarse_size(int(self.khmax_forcing / o.deltakx) + 1)
        ny = coarse_size(int(kymax / o.deltaky) + 1)
        if self.ndim == 2:
            return [nx, ny]
        nz = coarse_size(int(self.kvmax_forcing / o.deltakz) + 1)
        return [nx, ny, nz]

    def _compute_cond_no_forcing(self):
        oc = self.oper_c
        if self.anisotropic is None:
            K = oc.K
            C = np.logical_or(K > self.kmax_forcing, K < self.kmin_forcing)
        else:
            C = self._cond_anisotropic()
        if self.ndim == 2:
            nkyc, nkxc = oc.shapeK
            C[nkyc // 2, :] = True
            C[:, nkxc - 1] = True
        else:
            nkzc, nkyc, nkxc = oc.shapeK
            C[nkzc // 2, :, :] = True
            C[:, nkyc // 2, :] = True
            C[:, :, nkxc - 1] = True
        return C

    def _cond_anisotropic(self):
        """anisotropic.py:139-236."""
        oc = self.oper_c
        an = self.anisotropic
        angle = an["angle"]
        delta_angle = an.get("delta_angle")
        kz_neg = an.get("kz_negative_enable", False)
        if self.ndim == 2:
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
            return np.logical_or(c_kh, c_kv)
        K = oc.K.copy()
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
        return np.logical_or(c_k, c_t)

    # ------------------------------------------------------------ random
    def create_arrayK_random(self):
        """Uniform real and imaginary parts on [min_val, 1)."""
        shape = self.oper_c.shapeK
        lo = 0.0 if self.min_val is None else self.min_val
        re = np.random.random(shape)
        im = np.random.random(shape)
        return (lo + (1.0 - lo) * re) + 1j * (lo + (1.0 - lo) * im)

    def project_fft_on_realX(self, f):
        """operators2d.py:222-259 (2D), operators3d.py:1051-1052 (3D)."""
        oc = self.oper_c
        if self.ndim == 3:
            return oc.fft(oc.ifft(f))
        f = f.copy()
        nky = oc.shapeK[0]
        kyM = nky // 2
        kxM = oc.shapeK[1] - 1
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

    def compute_forcingc_raw(self):
        """RandomSimplePseudoSpectral.compute_forcingc_raw (754-762)."""
        f = self.create_arrayK_random()
        f[self.COND_NO_F] = 0.0
        return self.project_fft_on_realX(f)

    # ------------------------------------------------------------ per step
    def forcingc_raw_each_time(self, t):
        """specific.py:822-840 (new random array when the period ends)."""
        if t - self.t_last_change >= self.period_change_f0f1:
            self.t_last_change = t
            self._seed0 = self._seed1
            self.forcing0 = self.forcing1
            self._seed1 = np.random.randint(0, 2**31)
            np.random.seed(self._seed1)
            self.forcing1 = self.compute_forcingc_raw()
        omega = np.pi / self.period_change_f0f1
        deltaf = self.forcing1 - self.forcing0
        return self.forcing1 - 0.5 * (
            np.cos((t - self.t_last_change) * omega) + 1
        ) * deltaf

    def coarse_from_full(self, a_fft):
        """coarse_seq_from_fft_loc (operators2d.py:264-281, 3D alike)."""
        oc = self.oper_c
        idx = []
        for ax, nc in enumerate(oc.shapeK[:-1]):
            n = self.oper.shapeK[ax]
            idx.append(np.array([i if i <= nc / 2 else i - nc + n
                                 for i in range(nc)]))
        nkxc = oc.shapeK[-1]
        if self.ndim == 2:
            return a_fft[np.ix_(idx[0], np.arange(nkxc))].copy()
        return a_fft[np.ix_(idx[0], idx[1], np.arange(nkxc))].copy()

    def full_from_coarse(self, fc):
        """put_coarse_array_in_array_fft (operators2d.py:525-600)."""
        oc = self.oper_c
        out = np.zeros(self.oper.shapeK, dtype=np.complex128)
        idx = []
        for ax, nc in enumerate(oc.shapeK[:-1]):
            n = self.oper.shapeK[ax]
            idx.append(np.array([i if i <= nc / 2 else i - nc + n
                                 for i in range(nc)]))
        nkxc = oc.shapeK[-1]
        if self.ndim == 2:
            out[np.ix_(idx[0], np.arange(nkxc))] = fc
        else:
            out[np.ix_(idx[0], idx[1], np.arange(nkxc))] = fc
        return out

    def forced_variable(self, state_spect):
        """State value of the forced key (state.get_var / compute)."""
        if self.key_forced in self.keys:
            return state_spect[self.keys.index(self.key_forced)]
        if self.key_forced == "vt_fft":
            return self.oper.vtfft_from_vecfft(*state_spect[:3])
        raise ValueError(self.key_forced)

    def coarse_state_from_key(self, fc):
        """fstate_coarse.init_statespect_from(**{key: fc}) (base/state.py
        352-383; ns3d/state.py 146-169 for vt_fft)."""
        out = np.zeros((len(self.keys),) + self.oper_c.shapeK, np.complex128)
        if self.key_forced in self.keys:
            out[self.keys.index(self.key_forced)] = fc
        elif self.key_forced == "vt_fft":
            vx, vy, vz = self.oper_c.vecfft_from_vtfft(fc)
            out[0], out[1], out[2] = vx, vy, vz
        else:
            raise ValueError(self.key_forced)
        return out

    def compute(self, state_spect, t, dt):
        """NormalizedForcing.compute (specific.py:460-500) with the
        2nd-degree-equation normalisation (587-677)."""
        oc = self.oper_c
        a_fft = self.coarse_from_full(self.forced_variable(state_spect))
        fa = self.forcingc_raw_each_time(t)
        a = dt / 2 * oc.sum_wavenumbers(np.abs(fa) ** 2)
        b = oc.sum_wavenumbers((a_fft.conj() * fa).real)
        fa = fa * roots_2nd_degree(a, b, -self.forcing_rate, self.which_root)
        self.last_coarse = fa
        cstate = self.coarse_state_from_key(fa)
        oc.dealiasing(cstate)
        return np.array([self.full_from_coarse(c) for c in cstate])
