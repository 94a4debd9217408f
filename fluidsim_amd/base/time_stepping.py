"""Time stepping on the MI355X engine.

Mirrors TimeSteppingBase0 / TimeSteppingBase / TimeSteppingPseudoSpectral
(fluidsim/base/time_stepping/base.py:29-426,
fluidsim/base/time_stepping/pseudo_spect.py:155-984): same parameters, same
loop (start -> main_loop -> one_time_step), same integrating-factor RK4/RK2
algebra, CFL rule with 2 % hysteresis, NaN check raising ValueError, SIGUSR2
and max_elapsed stops.

Device dataflow of one RK4 step (all in HBM, no host round trip inside):
stage s: N = tendency(src)  -> fused inverse passes, products, forward passes
         rk_stage(s)        -> IF-RK algebra with E = exp(-dt sigma),
                               E2 = exp(-dt/2 sigma) in FP64 from indices,
                               projection (3D) folded in
With USE_CFL the stage-0 tendency also reduces max|u_i| (the velocity it
transforms anyway) and a one-thread kernel applies the CFL rule to the
device dt before stage 0's algebra, so dt follows base.py:320-366 from the
state at the start of the step, as in the reference.  One small device->host
read per step returns dt and the NaN flag.
"""

import signal
from datetime import datetime, timedelta
from time import time
from warnings import warn

import ctypes

from .. import _lib
from ..setofvariables import SetOfVariables


class TimeSteppingPseudoSpectral:
    @staticmethod
    def _complete_params_with_default(params):
        """base.py:33-94 + pseudo_spect.py:160-169."""
        attribs = {
            "USE_T_END": True,
            "t_end": 10.0,
            "it_end": 10,
            "USE_CFL": True,
            "type_time_scheme": "RK4",
            "deltat0": 0.2,
            "deltat_max": 0.2,
            "cfl_coef": None,
            "max_elapsed": None,
        }
        params._set_child("time_stepping", attribs=attribs)
        params.time_stepping._set_child(
            "phaseshift_random",
            attribs=dict(nb_pairs=1, nb_steps_compute_new_pair=None),
        )

    def __init__(self, sim):
        self.params = sim.params
        self.sim = sim
        self.it = 0
        self.t = 0
        self._stop_signal_received = False
        self._has_to_stop = False

        def handler_signals(signal_number, stack):
            print(f"signal {signal_number} received.")
            self._stop_signal_received = signal_number

        try:
            signal.signal(signal.SIGUSR2, handler_signals)
        except ValueError:
            warn("Cannot handle signals - is multithreading on?")
        except AttributeError:
            pass

        max_elapsed = getattr(self.params.time_stepping, "max_elapsed", None)
        if max_elapsed is not None:
            try:
                self.max_elapsed = float(max_elapsed)
            except ValueError:
                t = datetime.strptime(max_elapsed, "%H:%M:%S")
                self.max_elapsed = timedelta(
                    hours=t.hour, minutes=t.minute, seconds=t.second
                ).total_seconds()
            self._time_should_stop = time() + self.max_elapsed
        else:
            self.max_elapsed = None
        self.init_from_params()

    # ------------------------------------------------------------ init
    def init_from_params(self):
        self._init_compute_time_step()
        self._init_time_scheme()

    def _init_compute_time_step(self):
        """base.py:248-297."""
        p = self.params.time_stepping
        if p.USE_CFL:
            if p.cfl_coef is not None:
                self.CFL = p.cfl_coef
            elif p.type_time_scheme.startswith(("RK2", "Euler")):
                self.CFL = 0.4
            elif p.type_time_scheme.startswith("RK4"):
                self.CFL = 1.0
            else:
                raise ValueError("Problem name time_scheme")
        self.deltat = p.deltat0
        self.deltat_max = p.deltat_max
        self._dt_device = None
        if p.USE_CFL and not self.sim.state.has_vars("ux") and not self.sim.state.has_vars("vx"):
            raise ValueError("params_ts.USE_CFL but no velocity.")

    def _init_time_scheme(self):
        scheme = self.params.time_stepping.type_time_scheme
        if scheme == "RK4":
            self._scheme = 4
        elif scheme == "RK2":
            self._scheme = 2
        else:
            raise ValueError(f'Problem name time_scheme ("{scheme}")')
        like = self.sim.state.state_spect
        self._state_spect_tmp = SetOfVariables(like=like, info="tmp")
        self._state_spect_tmp1 = SetOfVariables(like=like, info="tmp1")
        self._tendencies = SetOfVariables(like=like, info="tendencies_nonlin")
        self._visc = _lib.doubles(self.sim.visc_coefs())
        self._scal = (ctypes.c_double * 6)()

    @property
    def freq_lin(self):
        f_d, f_d_hypo = self.sim.compute_freq_diss()
        return f_d + f_d_hypo

    # ------------------------------------------------------------ loop
    def start(self):
        self.main_loop(print_begin=True, save_init_field=True)
        self.finalize_main_loop()

    def prepare_main_loop(self):
        self.sim.__enter__()
        output = self.sim.output
        if not getattr(output, "_has_been_initialized_with_state", False):
            output.init_with_initialized_state()
        self._prepare_main_loop_called = True

    def finalize_main_loop(self):
        self.sim.__exit__()

    def main_loop(self, print_begin=False, save_init_field=False):
        if not hasattr(self, "_prepare_main_loop_called"):
            self.prepare_main_loop()
        print_stdout = self.sim.output.print_stdout
        if print_begin:
            print_stdout(
                "*************************************\nBeginning of the computation"
            )
        p = self.params.time_stepping
        if p.USE_T_END:
            print_stdout(f"    compute until t = {p.t_end:10.6g}")
            while self.t < p.t_end and not self._has_to_stop:
                self.one_time_step()
        else:
            print_stdout(f"    compute until it = {p.it_end:8d}")
            while self.it < p.it_end and not self._has_to_stop:
                self.one_time_step()

    def one_time_step(self):
        """base.py:208-244; the CFL update is fused into stage 0."""
        if self.max_elapsed is not None and time() > self._time_should_stop:
            self.sim.output.print_stdout(
                "Maximum elapsed time reached. Should stop soon."
            )
            self._has_to_stop = True
        if self._stop_signal_received:
            name = signal.Signals(self._stop_signal_received).name
            self.sim.output.print_stdout(
                f"Stop signal ({name}) received so _has_to_stop set to True"
            )
            self._has_to_stop = True
        if getattr(self.sim, "is_forcing_enabled", False):
            # base.py:212-213 (the normalisation itself runs on the device
            # inside the first stage, after the CFL update)
            self.sim.forcing.compute()
        self.sim.output.one_time_step()
        self.one_time_step_computation()
        self.t += self.deltat
        self.it += 1

    def is_simul_completed(self):
        p = self.params.time_stepping
        return self.t >= p.t_end if p.USE_T_END else self.it >= p.it_end

    # ------------------------------------------------------------ engine
    def _stream(self):
        return _lib.stream_ptr()

    def _push_dt(self):
        if self._dt_device != self.deltat:
            _lib.call("fsx_set_dt", self.sim.oper.plan.handle, float(self.deltat),
                      self._stream())
            self._dt_device = self.deltat

    def _read_scalars(self):
        _lib.call("fsx_get_scalars", self.sim.oper.plan.handle, self._scal,
                  self._stream())
        return list(self._scal)

    def compute_time_increment_CLF(self):
        """CFL rule of base.py:320-366 evaluated on the device from the
        current state (used for the first output; the loop fuses it)."""
        plan = self.sim.oper.plan.handle
        s = self._stream()
        self._push_dt()
        _lib.call("fsx_reset_maxima", plan, s)
        self.sim._engine_tendency(self.sim.state.state_spect.data,
                                  self._tendencies.data, want_maxima=True)
        _lib.call("fsx_cfl_update", plan, self.sim._cfl_ncomp, float(self.CFL),
                  float(self.deltat_max), s)
        sc = self._read_scalars()
        self.deltat = self._dt_device = sc[4]

    # fused stages: the tendency and the Runge-Kutta algebra of a stage in
    # one pass sequence (the tendency never goes to HBM); False selects the
    # separate tendency + stage kernels (same results, used to cross-check)
    fused = True

    def _time_step_RK(self):
        p = self.params.time_stepping
        sim = self.sim
        plan = sim.oper.plan.handle
        s = self._stream()
        S = sim.state.state_spect.data
        tmp = self._state_spect_tmp.data
        t1 = self._state_spect_tmp1.data
        self._push_dt()
        maker = (sim.forcing.forcing_maker
                 if getattr(sim, "is_forcing_enabled", False) else None)
        # the model term of a turbulence model needs the tendency array:
        # separate tendency + stage kernels then
        fused = self.fused and not (
            getattr(sim, "is_turb_model_enabled", False)
            and getattr(sim, "_turb_model_in_tendency", True)
        )
        if fused and hasattr(sim, "_engine_rk_stage"):
            cfl = float(self.CFL) if p.USE_CFL else 0.0
            for stage in range(self._scheme):
                sim._engine_rk_stage(self._scheme, stage, S, tmp, t1, self._visc,
                                     cfl if stage == 0 else 0.0,
                                     float(self.deltat_max))
            if maker is not None:
                maker._pending = False  # consumed by stage 0
            return
        N = self._tendencies.data
        nvar = S.shape[0]
        proj = sim._engine_projection
        novz = int(getattr(sim, "no_vz_kz0", False))
        if p.USE_CFL:
            _lib.call("fsx_reset_maxima", plan, s)
        sim._engine_tendency(S, N, want_maxima=bool(p.USE_CFL))
        if p.USE_CFL:
            _lib.call("fsx_cfl_update", plan, sim._cfl_ncomp, float(self.CFL),
                      float(self.deltat_max), s)
        if maker is not None and getattr(maker, "_pending", False):
            _lib.call("fsx_forcing_normalize", plan, _lib.dptr(S), s)
            maker._pending = False
        for stage in range(self._scheme):
            if stage > 0:
                sim._engine_tendency(t1, N, want_maxima=False)
            _lib.call("fsx_rk_stage", plan, self._scheme, stage, nvar,
                      _lib.dptr(S), _lib.dptr(tmp), _lib.dptr(t1), _lib.dptr(N),
                      proj, novz, self._visc, s)

    def one_time_step_computation(self):
        """pseudo_spect.py:236-243 / ns3d/time_stepping.py:8-20; projection
        and dealiasing of the new state are part of the last stage."""
        self._time_step_RK()
        sim = self.sim
        sim.state.invalidate_phys()
        sc = self._read_scalars()
        self.deltat = self._dt_device = sc[4]
        if sc[5] != 0:
            _lib.call("fsx_clear_nan_flag", sim.oper.plan.handle, self._stream())
            raise ValueError(f"nan at it = {self.it}, t = {self.t:.4f}")
