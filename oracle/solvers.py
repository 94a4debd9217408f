"""FP64 restatement of the ns2d / ns3d / ns3d.strat tendencies and of the
integrating-factor time stepping.

Test infrastructure only (see ``oracle/__init__.py``).
"""

import numpy as np

from .operators import Oper2D, Oper3D


def max_abs(arr):
    """fluidsim/base/time_stepping/base.py:25-26."""
    return max(abs(arr.min()), abs(arr.max()))


class _SolverBase:
    """State container + time stepping shared by the three solvers.

    state_spect is an ndarray (nvar, *shapeK) like fluidsim's SetOfVariables
    (fluidsim/base/state.py:243-248).
    """

    keys_state_spect = ()

    def __init__(
        self,
        oper,
        nu_2=0.0,
        nu_4=0.0,
        nu_8=0.0,
        nu_m4=0.0,
        type_time_scheme="RK4",
        USE_CFL=False,
        deltat0=0.2,
        deltat_max=0.2,
        cfl_coef=None,
    ):
        self.oper = oper
        self.nu_2, self.nu_4, self.nu_8, self.nu_m4 = nu_2, nu_4, nu_8, nu_m4
        self.type_time_scheme = type_time_scheme
        self.USE_CFL = USE_CFL
        self.deltat = float(deltat0)
        self.deltat_max = float(deltat_max)
        # fluidsim/base/time_stepping/base.py:252-266
        if cfl_coef is not None:
            self.CFL = cfl_coef
        elif type_time_scheme.startswith(("RK2", "Euler")):
            self.CFL = 0.4
        elif type_time_scheme.startswith("RK4"):
            self.CFL = 1.0
        else:
            raise ValueError("Problem name time_scheme")
        self.state_spect = np.zeros(
            (len(self.keys_state_spect),) + oper.shapeK, dtype=np.complex128
        )
        self.t = 0.0
        self.it = 0
        # forcing (oracle.forcing.TCRandomForcing or any object with
        # compute(state_spect, t, dt) -> full forcing), computed once per
        # step after the CFL update (base.py:208-214) and added to the
        # tendencies at every stage
        self.forcing = None
        self._F = None
        # turbulence model (oracle.turb_model.Smagorinsky) evaluated on the
        # state of each stage, added before the projection (3D)
        self.turb_model = None
        f_d, f_d_hypo = self.compute_freq_diss()
        self.freq_lin = f_d + f_d_hypo
        self._dt_coefs = None

    # --- linear operator ------------------------------------------------
    def compute_freq_diss(self):
        """fluidsim/base/solvers/pseudo_spect.py:134-191."""
        oper = self.oper
        if self.nu_2 > 0:
            f_d = self.nu_2 * oper.K2
        else:
            f_d = np.zeros_like(oper.K2)
        if self.nu_4 > 0.0:
            f_d = f_d + self.nu_4 * oper.K4
        if self.nu_8 > 0.0:
            f_d = f_d + self.nu_8 * oper.K8
        if self.nu_m4 != 0.0:
            f_d_hypo = self.nu_m4 / oper.K2_not0**2
            if f_d_hypo.ndim == 2:
                f_d_hypo[0, 0] = f_d_hypo[0, 1]
            else:
                f_d_hypo[0, 0, 0] = f_d_hypo[0, 0, 1]
        else:
            f_d_hypo = 0.0
        return f_d, f_d_hypo

    def exact_linear_coefs(self, dt):
        """ExactLinearCoefs.compute (pseudo_spect.py:122-141): E = exp(-dt s),
        E2 = exp(-dt/2 s), recomputed only when dt changes (143-148)."""
        if self._dt_coefs != dt:
            self._exact = np.exp(-dt * self.freq_lin)
            self._exact2 = np.exp(-dt / 2 * self.freq_lin)
            self._dt_coefs = dt
        return self._exact, self._exact2

    # --- CFL ------------------------------------------------------------
    def cfl_tmp(self):
        raise NotImplementedError

    def compute_time_increment_CLF(self):
        """_compute_time_increment_CLF_from_tmp (base.py:341-354)."""
        tmp = self.cfl_tmp()
        if tmp > 0:
            deltat_CFL = self.CFL / tmp
        else:
            deltat_CFL = self.deltat_max
        maybe_new_dt = min(deltat_CFL, self.deltat_max)
        normalize_diff = abs(self.deltat - maybe_new_dt) / maybe_new_dt
        if normalize_diff > 0.02:
            self.deltat = maybe_new_dt

    # --- time schemes -----------------------------------------------------
    def project_state_spect(self, state_spect):
        pass

    def _time_step_RK4(self):
        """pseudo_spect.py:798-984 (rk4_step0 .. rk4_step3)."""
        dt = self.deltat
        diss, diss2 = self.exact_linear_coefs(dt)
        S = self.state_spect
        N0 = self.tendencies_nonlin(S)
        tmp = (S + dt / 6 * N0) * diss
        t1 = (S + dt / 2 * N0) * diss2
        N1 = self.tendencies_nonlin(t1)
        tmp += dt / 3 * diss2 * N1
        t1 = S * diss2 + dt / 2 * N1
        N2 = self.tendencies_nonlin(t1)
        tmp += dt / 3 * diss2 * N2
        t1 = S * diss + dt * diss2 * N2
        N3 = self.tendencies_nonlin(t1)
        S[...] = tmp + dt / 6 * N3

    def _time_step_RK2(self):
        """pseudo_spect.py:469-517: midpoint with exp factors."""
        dt = self.deltat
        diss, diss2 = self.exact_linear_coefs(dt)
        S = self.state_spect
        N0 = self.tendencies_nonlin(S)
        t12 = (S + dt / 2 * N0) * diss2
        N1 = self.tendencies_nonlin(t12)
        S[...] = S * diss + dt * diss2 * N1

    # --- Euler, RK2 trapezoid and phase-shifted schemes (pseudo_spect.py
    # :245-796); random variants take their translations from
    # self.phase_pairs (a callable returning two shift vectors per step)
    def phaseshift(self, shift):
        """exp(i shift . k) on the K grid (_get_phaseshift, 281-300)."""
        o = self.oper
        if self.state_spect.ndim == 3:
            phase = shift[0] * o.KX + shift[1] * o.KY
        else:
            phase = shift[0] * o.Kx + shift[1] * o.Ky + shift[2] * o.Kz
        return np.exp(1j * phase)

    def half_shift(self):
        o = self.oper
        return (0.5 * o.deltax, 0.5 * o.deltay, 0.5 * getattr(o, "deltaz", 0.0))

    def _time_step_Euler(self):
        """pseudo_spect.py:245-272."""
        dt = self.deltat
        E, _ = self.exact_linear_coefs(dt)
        S = self.state_spect
        N0 = self.tendencies_nonlin(S)
        S[...] = (S + dt * N0) * E

    def _time_step_Euler_phaseshift(self):
        """pseudo_spect.py:374-419."""
        dt = self.deltat
        E, _ = self.exact_linear_coefs(dt)
        S = self.state_spect
        ps = self.phaseshift(self.half_shift())
        N0 = self.tendencies_nonlin(S)
        Ns = self.tendencies_nonlin(ps * S) / ps
        S[...] = (S + dt * 0.5 * (N0 + Ns)) * E

    def _time_step_Euler_phaseshift_random(self):
        """pseudo_spect.py:421-467."""
        dt = self.deltat
        E, _ = self.exact_linear_coefs(dt)
        S = self.state_spect
        sa, sb = self.phase_pairs()
        pa, pb = self.phaseshift(sa), self.phaseshift(sb)
        Na = self.tendencies_nonlin(pa * S) / pa
        Nb = self.tendencies_nonlin(pb * S) / pb
        S[...] = (S + dt * 0.5 * (Na + Nb)) * E

    def _time_step_RK2_trapezoid(self):
        """pseudo_spect.py:519-569 (Heun)."""
        dt = self.deltat
        E, _ = self.exact_linear_coefs(dt)
        S = self.state_spect
        N0 = self.tendencies_nonlin(S)
        S1 = (S + dt * N0) * E
        N1 = self.tendencies_nonlin(S1)
        S[...] = (S + dt / 2 * N0) * E + dt / 2 * N1

    def _time_step_RK2_phaseshift(self):
        """pseudo_spect.py:571-639."""
        dt = self.deltat
        E, E2 = self.exact_linear_coefs(dt)
        S = self.state_spect
        N0 = self.tendencies_nonlin(S)
        S1 = (S + dt * N0) * E
        ps = self.phaseshift(self.half_shift())
        N1s = self.tendencies_nonlin(ps * S1)
        Nd = 0.5 * (N0 + N1s / ps)
        S[...] = S * E + dt * E2 * Nd

    def _time_step_RK2_phaseshift_random(self):
        """pseudo_spect.py:641-710."""
        dt = self.deltat
        E, E2 = self.exact_linear_coefs(dt)
        S = self.state_spect
        sa, sb = self.phase_pairs()
        pa, pb = self.phaseshift(sa), self.phaseshift(sb)
        N0 = self.tendencies_nonlin(pa * S) / pa
        S1 = (S + dt * N0) * E
        N1s = self.tendencies_nonlin(pb * S1)
        Nd = 0.5 * (N0 + N1s / pb)
        S[...] = S * E + dt * E2 * Nd

    def _time_step_RK2_phaseshift_exact(self):
        """pseudo_spect.py:712-796."""
        dt = self.deltat
        E, E2 = self.exact_linear_coefs(dt)
        S = self.state_spect
        ps = self.phaseshift(self.half_shift())
        N0 = self.tendencies_nonlin(S)
        N0s = self.tendencies_nonlin(ps * S)
        Nd0 = 0.5 * (N0 + N0s / ps)
        S1 = (S + dt * Nd0) * E
        N1 = self.tendencies_nonlin(S1)
        N1s = self.tendencies_nonlin(ps * S1)
        Nd = 0.5 * (Nd0 + 0.5 * (N1 + N1s / ps))
        S[...] = S * E + dt * E2 * Nd

    def one_time_step_computation(self):
        """pseudo_spect.py:236-243 / ns3d/time_stepping.py:8-20."""
        step = getattr(self, "_time_step_" + self.type_time_scheme, None)
        if step is None:
            raise ValueError(self.type_time_scheme)
        step()
        self.project_state_spect(self.state_spect)
        self.oper.dealiasing(self.state_spect)
        if np.isnan(np.sum(self.state_spect[0])):
            raise ValueError(f"nan at it = {self.it}, t = {self.t:.4f}")

    def one_time_step(self):
        """base.py:208-244 (CFL first, forcing, then the step, t, it)."""
        if self.USE_CFL:
            self.compute_time_increment_CLF()
        if self.forcing is not None:
            self._F = self.forcing.compute(self.state_spect, self.t, self.deltat)
        self.one_time_step_computation()
        self.t += self.deltat
        self.it += 1

    def run(self, nsteps, record_dt=False):
        dts = []
        for _ in range(nsteps):
            self.one_time_step()
            dts.append(self.deltat)
        return dts if record_dt else None

    def run_with_dts(self, dts):
        """Advance with a prescribed dt sequence (no CFL)."""
        for dt in dts:
            self.deltat = float(dt)
            self.one_time_step_computation()
            self.t += self.deltat
            self.it += 1


class NS2D(_SolverBase):
    """fluidsim/solvers/ns2d/solver.py:111-194, state.py:95-119."""

    keys_state_spect = ("rot_fft",)

    def __init__(self, oper: Oper2D, beta=0.0, **kwargs):
        self.beta = beta
        super().__init__(oper, **kwargs)

    def tendencies_nonlin(self, state_spect):
        oper = self.oper
        rot_fft = state_spect[0]
        ux_fft, uy_fft = oper.vecfft_from_rotfft(rot_fft)
        ux = oper.ifft(ux_fft)
        uy = oper.ifft(uy_fft)
        px_rot_fft, py_rot_fft = oper.gradfft_from_fft(rot_fft)
        px_rot = oper.ifft(px_rot_fft)
        py_rot = oper.ifft(py_rot_fft)
        # compute_Frot (solver.py:33-39)
        if self.beta == 0:
            Frot = -ux * px_rot - uy * py_rot
        else:
            Frot = -ux * px_rot - uy * (py_rot + self.beta)
        out = np.empty_like(state_spect)
        out[0] = oper.fft(Frot)
        oper.dealiasing(out[0])
        if self._F is not None:
            out += self._F  # solver.py:190-191, after the dealiasing
        return out

    def init_from_rotfft(self, rot_fft):
        """state.py:115-119: dealias then set."""
        rot_fft = np.array(rot_fft, dtype=np.complex128)
        self.oper.dealiasing(rot_fft)
        self.state_spect[0] = rot_fft

    def state_phys(self):
        """statephys_from_statespect (state.py:95-106): rot, ux, uy."""
        oper = self.oper
        rot_fft = self.state_spect[0]
        ux_fft, uy_fft = oper.vecfft_from_rotfft(rot_fft)
        return {
            "rot": oper.ifft(rot_fft),
            "ux": oper.ifft(ux_fft),
            "uy": oper.ifft(uy_fft),
        }

    def cfl_tmp(self):
        """_compute_time_increment_CLF_uxuy (base.py:356-366)."""
        ph = self.state_phys()
        return max_abs(ph["ux"]) / self.oper.deltax + max_abs(
            ph["uy"]
        ) / self.oper.deltay

    def compute_energy_fft(self):
        """ns2d/output/__init__.py:110-114."""
        ux_fft, uy_fft = self.oper.vecfft_from_rotfft(self.state_spect[0])
        return (np.abs(ux_fft) ** 2 + np.abs(uy_fft) ** 2) / 2.0

    def compute_enstrophy_fft(self):
        return np.abs(self.state_spect[0]) ** 2 / 2.0

    def compute_energy(self):
        return self.oper.sum_wavenumbers(self.compute_energy_fft())

    def compute_enstrophy(self):
        return self.oper.sum_wavenumbers(self.compute_enstrophy_fft())

    def spatial_means(self):
        """SpatialMeansNS2D._save_one_time without forcing
        (ns2d/output/spatial_means.py:23-38)."""
        e_fft = self.compute_energy_fft()
        z_fft = self.compute_enstrophy_fft()
        f_d, f_d_hypo = self.compute_freq_diss()
        sw = self.oper.sum_wavenumbers
        return {
            "E": sw(e_fft),
            "Z": sw(z_fft),
            "epsK": sw(f_d * 2 * e_fft),
            "epsK_hypo": sw(f_d_hypo * 2 * e_fft),
            "epsZ": sw(f_d * 2 * z_fft),
            "epsZ_hypo": sw(f_d_hypo * 2 * z_fft),
        }


class NS3D(_SolverBase):
    """fluidsim/solvers/ns3d/solver.py:147-263."""

    keys_state_spect = ("vx_fft", "vy_fft", "vz_fft")

    def __init__(
        self, oper: Oper3D, f=None, projection=None, no_vz_kz0=False, **kwargs
    ):
        self.f = f
        self.no_vz_kz0 = no_vz_kz0
        # _init_projection (solver.py:147-174)
        if projection is None:
            self._projector = oper.project_perpk3d
        elif projection in ("toroidal", "vortical"):
            self._projector = oper.project_toroidal
        elif projection == "poloidal":
            self._projector = oper.project_poloidal
        else:
            raise ValueError(projection)
        if no_vz_kz0:
            self.where_kz_0 = np.abs(oper.Kz) == 0.0
        super().__init__(oper, **kwargs)

    def project_state_spect(self, state_spect):
        """solver.py:255-263."""
        self._projector(state_spect[0], state_spect[1], state_spect[2])
        if self.no_vz_kz0:
            state_spect[2][self.where_kz_0] = 0.0
            if state_spect.shape[0] > 3:
                state_spect[3][self.where_kz_0] = 0.0

    def _omega_v_phys(self, state_spect):
        oper = self.oper
        vx_fft, vy_fft, vz_fft = state_spect[0], state_spect[1], state_spect[2]
        ox_fft, oy_fft, oz_fft = oper.rotfft_from_vecfft(vx_fft, vy_fft, vz_fft)
        if self.f is not None:
            oz_fft.flat[0] += self.f  # _modif_omegafft_with_f (176-178)
        omega = [oper.ifft(a) for a in (ox_fft, oy_fft, oz_fft)]
        v = [oper.ifft(a) for a in (vx_fft, vy_fft, vz_fft)]
        return v, omega

    def tendencies_nonlin(self, state_spect):
        """solver.py:180-253 (no forcing / turb_model)."""
        oper = self.oper
        v, omega = self._omega_v_phys(state_spect)
        fx, fy, fz = oper.vector_product(*v, *omega)
        out = np.empty_like(state_spect)
        out[0] = oper.fft(fx)
        out[1] = oper.fft(fy)
        out[2] = oper.fft(fz)
        if self._F is not None:
            out += self._F  # solver.py:243-244, before the projection
        if self.turb_model is not None:  # solver.py:246-249
            fxm, fym, fzm = self.turb_model.get_forcing(*state_spect[:3])
            out[0] += fxm
            out[1] += fym
            out[2] += fzm
        self.project_state_spect(out)
        oper.dealiasing(out)
        return out

    def init_from_vxvyvzfft(self, vx_fft, vy_fft, vz_fft):
        """ns3d/state.py:126-132 (ifft/fft round trip) followed by the
        projection done by InitFieldsNS3D (init_fields.py:210-212 via
        init_state_from_fieldsfft)."""
        oper = self.oper
        for i, a in enumerate((vx_fft, vy_fft, vz_fft)):
            self.state_spect[i] = oper.fft(oper.ifft(a))

    def state_phys(self):
        return {
            k[:-4]: self.oper.ifft(self.state_spect[i])
            for i, k in enumerate(self.keys_state_spect)
        }

    def cfl_tmp(self):
        """_compute_time_increment_CLF_uxuyuz (base.py:320-339)."""
        ph = self.state_phys()
        oper = self.oper
        return (
            max_abs(ph["vx"]) / oper.deltax
            + max_abs(ph["vy"]) / oper.deltay
            + max_abs(ph["vz"]) / oper.deltaz
        )

    def compute_energy_fft(self):
        """ns3d/output/__init__.py:101-112."""
        S = self.state_spect
        return 0.5 * (np.abs(S[0]) ** 2 + np.abs(S[1]) ** 2 + np.abs(S[2]) ** 2)

    def compute_energy(self):
        return self.oper.sum_wavenumbers(self.compute_energy_fft())

    def spatial_means(self):
        """Energy and dissipation (ns3d/output/spatial_means.py:23-73)."""
        e_fft = self.compute_energy_fft()
        f_d, f_d_hypo = self.compute_freq_diss()
        sw = self.oper.sum_wavenumbers
        return {
            "E": sw(e_fft),
            "epsK": sw(f_d * 2 * e_fft),
            "epsK_hypo": sw(f_d_hypo * 2 * e_fft),
        }


class NS3DStrat(NS3D):
    """fluidsim/solvers/ns3d/strat/solver.py:29-33, 138-216."""

    keys_state_spect = ("vx_fft", "vy_fft", "vz_fft", "b_fft")

    def __init__(self, oper: Oper3D, N=1.0, **kwargs):
        self.N = N
        super().__init__(oper, **kwargs)

    def tendencies_nonlin(self, state_spect):
        oper = self.oper
        v, omega = self._omega_v_phys(state_spect)
        fx, fy, fz = oper.vector_product(*v, *omega)
        out = np.empty_like(state_spect)
        out[0] = oper.fft(fx)
        out[1] = oper.fft(fy)
        out[2] = oper.fft(fz)
        out[2] += state_spect[3]  # fz_fft += b_fft (198)
        b = oper.ifft(state_spect[3])
        div_vb_fft = oper.div_vb_fft_from_vb(v[0], v[1], v[2], b)
        # compute_fb_fft (29-33)
        out[3] = -div_vb_fft - self.N**2 * state_spect[2]
        if self._F is not None:
            out += self._F  # strat/solver.py:209-213
        self.project_state_spect(out)
        oper.dealiasing(out)
        return out

    def init_from_fft(self, vx_fft, vy_fft, vz_fft, b_fft):
        oper = self.oper
        for i, a in enumerate((vx_fft, vy_fft, vz_fft, b_fft)):
            self.state_spect[i] = oper.fft(oper.ifft(a))

    def compute_energies(self):
        """E_K and E_A = |b|^2 / (2 N^2) (strat/output; test_solver.py:35-58)."""
        S = self.state_spect
        sw = self.oper.sum_wavenumbers
        EK = sw(0.5 * (np.abs(S[0]) ** 2 + np.abs(S[1]) ** 2 + np.abs(S[2]) ** 2))
        EA = sw(0.5 * np.abs(S[3]) ** 2 / self.N**2)
        return EK, EA
