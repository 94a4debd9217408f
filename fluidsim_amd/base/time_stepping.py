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
            # added (no reference counterpart): replay the fused Runge-Kutta
            # step from a HIP graph
            "hip_graph": False,
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

    # schemes built from full tendencies and IF linear combinations
    # (pseudo_spect.py:245-796)
    _GENERIC_SCHEMES = (
        "Euler", "Euler_phaseshift", "Euler_phaseshift_random", "RK2_trapezoid",
        "RK2_phaseshift", "RK2_phaseshift_random", "RK2_phaseshift_exact",
    )

    def _init_time_scheme(self):
        scheme = self.params.time_stepping.type_time_scheme
        like = self.sim.state.state_spect
        if scheme == "RK4":
            self._scheme = 4
        elif scheme == "RK2":
            self._scheme = 2
        elif scheme in self._GENERIC_SCHEMES:
            self._scheme = 0
            self._extra = [SetOfVariables(like=like, info=f"tmp{i}") for i in range(4)]
            if scheme.endswith("_random"):
                self._init_phaseshift_random()
        else:
            raise ValueError(f'Problem name time_scheme ("{scheme}")')
        self._state_spect_tmp = SetOfVariables(like=like, info="tmp")
        self._state_spect_tmp1 = SetOfVariables(like=like, info="tmp1")
        self._tendencies = SetOfVariables(like=like, info="tendencies_nonlin")
        self._visc = _lib.doubles(self.sim.visc_coefs())
        self._scal = (ctypes.c_double * 6)()

    # ------------------------------------------------------------ phase shifts
    def _shift_half(self):
        """_get_phaseshift (pseudo_spect.py:281-300): a translation by half
        a grid step along every axis, phase 0.5 (dx kx + dy ky + dz kz)."""
        o = self.sim.oper
        return (0.5 * o.deltax, 0.5 * o.deltay, 0.5 * getattr(o, "deltaz", 0.0))

    def _random_pair(self):
        """get_phases_random (operators2d.py:618-631, operators3d.py
        :1128-1146): translations alpha dx and beta dx with alpha uniform
        on [-1/2, 1/2) and |alpha - beta| = 1/2, per axis."""
        o = self.sim.oper
        deltas = (o.deltax, o.deltay) + ((o.deltaz,) if o.dim == 3 else ())
        alphas = [self._uniform(-0.5, 0.5) for _ in deltas]
        betas = [a + 0.5 if a < 0 else a - 0.5 for a in alphas]
        sa = tuple(a * d for a, d in zip(alphas, deltas)) + (0.0,) * (3 - len(deltas))
        sb = tuple(b * d for b, d in zip(betas, deltas)) + (0.0,) * (3 - len(deltas))
        return sa, sb

    def _init_phaseshift_random(self):
        """pseudo_spect.py:302-327; the translations come from Python's
        random module as in operators2d.py:13,618-631, so random.seed
        reproduces a run."""
        import random

        self._uniform = random.uniform
        p = self.params.time_stepping.phaseshift_random
        if p.nb_steps_compute_new_pair is None:
            p.nb_steps_compute_new_pair = 2 if p.nb_pairs == 1 else 4 * p.nb_pairs
        self._index_phaseshift = 1
        self._previous_index_pair = 0
        self._previous_index_flip = 0
        self._pairs_phaseshift = [self._random_pair() for _ in range(p.nb_pairs)]

    def _get_phaseshift_random(self):
        """pseudo_spect.py:329-371 (the pair rotation and renewal)."""
        from random import randint

        p = self.params.time_stepping.phaseshift_random
        nb_pairs, nb_new = p.nb_pairs, p.nb_steps_compute_new_pair
        if nb_pairs == 1 and nb_new == 1:
            alpha, beta = self._pairs_phaseshift[0]
        elif nb_pairs == 1 and nb_new == 2:
            pair = self._pairs_phaseshift[0]
            alpha, beta = pair if self._index_phaseshift == 1 else pair[::-1]
        else:
            index_pair = randint(0, nb_pairs - 1)
            pair = self._pairs_phaseshift[index_pair]
            index_flip = randint(0, 1)
            if (index_pair == self._previous_index_pair
                    and index_flip == self._previous_index_flip):
                index_flip = 0 if index_flip else 1
            self._previous_index_pair = index_pair
            self._previous_index_flip = index_flip
            alpha, beta = pair if index_flip else pair[::-1]
        if self._index_phaseshift == nb_new:
            self._index_phaseshift = 1
            self._pairs_phaseshift.pop(0)
            self._pairs_phaseshift.append(self._random_pair())
        else:
            self._index_phaseshift += 1
        return alpha, beta

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
        if self.sim.output.outputs_due():
            # outputs at t read dt and the forcing of this step: update both
            # first, as base.py:208-244 orders them (the step then skips them)
            self._update_dt_and_forcing()
        self.sim.output.one_time_step()
        self.one_time_step_computation()
        self.t += self.deltat
        self.it += 1

    # set by _update_dt_and_forcing: the coming step's CFL update and forcing
    # normalisation are done
    _pre_done = False

    def _update_dt_and_forcing(self):
        """CFL update (base.py:320-366) and forcing normalisation of the coming
        step on the current state, ahead of the outputs; the same device
        work the first stage of the step would otherwise do, so the step is
        unchanged."""
        p = self.params.time_stepping
        sim = self.sim
        maker = (sim.forcing.forcing_maker
                 if getattr(sim, "is_forcing_enabled", False) else None)
        pending = maker is not None and getattr(maker, "_pending", False)
        if not (p.USE_CFL or pending):
            return
        if p.USE_CFL:
            self.compute_time_increment_CLF()
        if pending:
            self._push_dt()
            if getattr(self, "_forcing_graph", False):
                _lib.call("fsx_forcing_push", sim.oper.plan.handle, self._stream())
            _lib.call("fsx_forcing_normalize", sim.oper.plan.handle,
                      _lib.dptr(sim.state.state_spect.data), self._stream())
            maker._pending = False
        self._pre_done = True

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
        use_cfl = p.USE_CFL and not self._pre_done
        if fused and hasattr(sim, "_engine_rk_stage"):
            cfl = float(self.CFL) if use_cfl else 0.0
            if getattr(p, "hip_graph", False) and not self._pre_done:
                self._graph_step(cfl, maker)
            else:
                self._issue_stages(cfl)
            if maker is not None:
                maker._pending = False  # consumed by stage 0
            return
        N = self._tendencies.data
        nvar = S.shape[0]
        proj = sim._engine_projection
        novz = int(getattr(sim, "no_vz_kz0", False))
        if use_cfl:
            _lib.call("fsx_reset_maxima", plan, s)
        sim._engine_tendency(S, N, want_maxima=bool(use_cfl))
        if use_cfl:
            _lib.call("fsx_cfl_update", plan, sim._cfl_ncomp, float(self.CFL),
                      float(self.deltat_max), s)
        if maker is not None and getattr(maker, "_pending", False):
            if getattr(self, "_forcing_graph", False):
                _lib.call("fsx_forcing_push", plan, s)
            _lib.call("fsx_forcing_normalize", plan, _lib.dptr(S), s)
            maker._pending = False
        for stage in range(self._scheme):
            if stage > 0:
                sim._engine_tendency(t1, N, want_maxima=False)
            _lib.call("fsx_rk_stage", plan, self._scheme, stage, nvar,
                      _lib.dptr(S), _lib.dptr(tmp), _lib.dptr(t1), _lib.dptr(N),
                      proj, novz, self._visc, s)

    def _issue_stages(self, cfl):
        sim = self.sim
        S = sim.state.state_spect.data
        tmp = self._state_spect_tmp.data
        t1 = self._state_spect_tmp1.data
        for stage in range(self._scheme):
            sim._engine_rk_stage(self._scheme, stage, S, tmp, t1, self._visc,
                                 cfl if stage == 0 else 0.0, float(self.deltat_max))

    _gstream = None

    def _graph_step(self, cfl, maker):
        """The fused step replayed from a HIP graph (params.time_stepping
        .hip_graph): the launches of the stages are captured once per
        variant (CFL update in stage 0 or not, forcing, the solver's own key)
        on a stream of the time stepping and replayed on it.  Everything a
        step changes lives in device memory: dt (CFL rule), the state and
        scratch arrays, the forcing box and, in graph mode, the forcing
        normalisation's coefficients, pushed before each replay.  The first
        step of a variant runs eagerly (kernel attributes, buffers the
        stages allocate), the second is captured, then replayed."""
        import torch

        sim = self.sim
        plan = sim.oper.plan.handle
        if self._gstream is None:
            self._gstream = torch.cuda.Stream()
            self._graphs = {}
        if maker is not None and not getattr(self, "_forcing_graph", False):
            _lib.call("fsx_forcing_graph", plan, 1)
            self._forcing_graph = True
        key = (cfl > 0, maker is not None,
               sim._graph_key() if hasattr(sim, "_graph_key") else None)
        cur = torch.cuda.current_stream()
        gs = self._gstream
        gs.wait_stream(cur)
        with torch.cuda.stream(gs):
            sp = _lib.stream_ptr(gs)
            if maker is not None:
                _lib.call("fsx_forcing_push", plan, sp)
            g = self._graphs.get(key)
            if g is None:
                self._issue_stages(cfl)
                self._graphs[key] = "eager"
            elif g == "eager":
                h = ctypes.c_void_p()
                _lib.call("fsx_graph_begin", sp)
                try:
                    self._issue_stages(cfl)
                finally:
                    rc = _lib.get_lib().fsx_graph_end(sp, ctypes.byref(h))
                _lib.check(rc, "fsx_graph_end")
                self._graphs[key] = h
                _lib.call("fsx_graph_launch", h, sp)
            else:
                _lib.call("fsx_graph_launch", g, sp)
        cur.wait_stream(gs)

    def __del__(self):
        for g in getattr(self, "_graphs", {}).values():
            if isinstance(g, ctypes.c_void_p) and g.value:
                try:
                    _lib.call("fsx_graph_destroy", g)
                except Exception:
                    pass

    # ------------------------------------------------------------ generic
    def _combine(self, out, terms, shift=(0.0, 0.0, 0.0), check_nan=False):
        """out = sum (a + b dt) exp(-e dt sigma) exp(i p shift.k) x over the
        terms (x, a, b, e, p) (fsx_if_combine)."""
        n = len(terms)
        ins = (ctypes.c_void_p * n)(*[_lib.dptr(t[0].data) for t in terms])
        a = _lib.doubles([t[1] for t in terms])
        b = _lib.doubles([t[2] for t in terms])
        e = _lib.doubles([t[3] for t in terms])
        pw = (ctypes.c_int * n)(*[t[4] for t in terms])
        _lib.call("fsx_if_combine", self.sim.oper.plan.handle, out.nvar,
                  _lib.dptr(out.data), n, ins, a, b, e, pw, _lib.doubles(shift),
                  self._visc, int(check_nan), self._stream())

    def _tend(self, x, out):
        self.sim.tendencies_nonlin(x, old=out)
        return out

    def _tend_shifted(self, x, shift, work, out):
        """N(ps x) (the caller divides by ps in its combination)."""
        self._combine(work, [(x, 1.0, 0.0, 0.0, 1)], shift)
        return self._tend(work, out)

    def _time_step_generic(self):
        """Euler, RK2_trapezoid and the phase-shifted schemes
        (pseudo_spect.py:245-796) from full tendencies and IF linear
        combinations; the state S is updated in place."""
        p = self.params.time_stepping
        sim = self.sim
        plan = sim.oper.plan.handle
        s = self._stream()
        self._push_dt()
        if p.USE_CFL and not self._pre_done:
            self.compute_time_increment_CLF()
        maker = (sim.forcing.forcing_maker
                 if getattr(sim, "is_forcing_enabled", False) else None)
        if maker is not None and getattr(maker, "_pending", False):
            if getattr(self, "_forcing_graph", False):
                _lib.call("fsx_forcing_push", plan, s)
            _lib.call("fsx_forcing_normalize", plan, _lib.dptr(sim.state.state_spect.data), s)
            maker._pending = False
        S = sim.state.state_spect
        T0, T1, T2, T3 = self._extra
        W = self._state_spect_tmp1
        scheme = p.type_time_scheme
        h = self._shift_half()
        if scheme == "Euler":
            N0 = self._tend(S, T0)
            self._combine(S, [(S, 1, 0, 1, 0), (N0, 0, 1, 1, 0)], check_nan=True)
        elif scheme == "Euler_phaseshift":
            N0 = self._tend(S, T0)
            N0s = self._tend_shifted(S, h, W, T1)
            self._combine(S, [(S, 1, 0, 1, 0), (N0, 0, 0.5, 1, 0), (N0s, 0, 0.5, 1, -1)],
                          h, check_nan=True)
        elif scheme == "Euler_phaseshift_random":
            sa, sb = self._get_phaseshift_random()
            Na = self._tend_shifted(S, sa, W, T0)
            Nb = self._tend_shifted(S, sb, W, T1)
            self._combine(T2, [(S, 1, 0, 1, 0), (Na, 0, 0.5, 1, -1)], sa)
            self._combine(S, [(T2, 1, 0, 0, 0), (Nb, 0, 0.5, 1, -1)], sb, check_nan=True)
        elif scheme == "RK2_trapezoid":
            N0 = self._tend(S, T0)
            self._combine(T1, [(S, 1, 0, 1, 0), (N0, 0, 1, 1, 0)])
            N1 = self._tend(T1, T2)
            self._combine(S, [(S, 1, 0, 1, 0), (N0, 0, 0.5, 1, 0), (N1, 0, 0.5, 0, 0)],
                          check_nan=True)
        elif scheme == "RK2_phaseshift":
            N0 = self._tend(S, T0)
            self._combine(T1, [(S, 1, 0, 1, 0), (N0, 0, 1, 1, 0)])
            N1s = self._tend_shifted(T1, h, W, T2)
            self._combine(S, [(S, 1, 0, 1, 0), (N0, 0, 0.5, 0.5, 0),
                              (N1s, 0, 0.5, 0.5, -1)], h, check_nan=True)
        elif scheme == "RK2_phaseshift_random":
            sa, sb = self._get_phaseshift_random()
            N0s = self._tend_shifted(S, sa, W, T0)
            # N0 = N0s / ps_alpha
            self._combine(T3, [(N0s, 1, 0, 0, -1)], sa)
            self._combine(T1, [(S, 1, 0, 1, 0), (T3, 0, 1, 1, 0)])
            N1s = self._tend_shifted(T1, sb, W, T2)
            self._combine(S, [(S, 1, 0, 1, 0), (T3, 0, 0.5, 0.5, 0),
                              (N1s, 0, 0.5, 0.5, -1)], sb, check_nan=True)
        elif scheme == "RK2_phaseshift_exact":
            N0 = self._tend(S, T0)
            N0s = self._tend_shifted(S, h, W, T1)
            # Nd0 = (N0 + N0s / ps) / 2 into T3
            self._combine(T3, [(N0, 0.5, 0, 0, 0), (N0s, 0.5, 0, 0, -1)], h)
            self._combine(T1, [(S, 1, 0, 1, 0), (T3, 0, 1, 1, 0)])
            N1 = self._tend(T1, T0)
            N1s = self._tend_shifted(T1, h, W, T2)
            # S E + dt E2 (Nd0 / 2 + N1 / 4 + N1s / (4 ps))
            self._combine(S, [(S, 1, 0, 1, 0), (T3, 0, 0.5, 0.5, 0), (N1, 0, 0.25, 0.5, 0),
                              (N1s, 0, 0.25, 0.5, -1)], h, check_nan=True)
        else:
            raise ValueError(scheme)
        # one_time_step_computation: projection (ns3d) and dealiasing
        if hasattr(sim, "project_state_spect") and sim.oper.dim == 3:
            sim.project_state_spect(S)
        sim.oper.dealiasing(S)

    def one_time_step_computation(self):
        """pseudo_spect.py:236-243 / ns3d/time_stepping.py:8-20; projection
        and dealiasing of the new state are part of the last stage."""
        try:
            if self._scheme == 0:
                self._time_step_generic()
            else:
                self._time_step_RK()
        finally:
            self._pre_done = False
        sim = self.sim
        sim.state.invalidate_phys()
        sc = self._read_scalars()
        self.deltat = self._dt_device = sc[4]
        if sc[5] != 0:
            _lib.call("fsx_clear_nan_flag", sim.oper.plan.handle, self._stream())
            raise ValueError(f"nan at it = {self.it}, t = {self.t:.4f}")
