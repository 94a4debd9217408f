"""Ensembles of independent ns2d runs advanced together on one GPU.

fluidsim runs one ``Simul`` per process group; an ensemble of many small
runs is new here.  Each run is an ns2d simulation (vorticity form, IF-RK4 or
RK2 at a fixed dt of its own, its own viscosities), the same algorithm as
``fluidsim_amd.solvers.ns2d`` (fluidsim/solvers/ns2d/solver.py:111-194,
fluidsim/base/time_stepping/pseudo_spect.py:466-517, 798-984).

Runs of one grid size form a group: one batch plan
(``fsx_plan_create_batch``) whose state arrays hold the group's runs one
after the other, (nb, n, n/2 + 1), and whose passes take the run as their
outer index, so a Runge-Kutta stage of every run of the group is three
launches.  A step of the whole ensemble is captured once into a HIP graph
and replayed (``fsx_graph_*``): the groups are parallel branches of the
graph (each group's stages on a stream of its own, forked from and joined
back to the capturing stream by events), so the launches of different
sizes overlap on the GPU.  A small autotuner picks, per group, the lines
per tile of the batched y-passes and the row pass (register rows with
radix-8 or radix-4 stages, or the LDS Stockham row pass) by timing every
pair on the group itself, and caches the choice in a JSON file keyed by
(size, batch, precision, scheme, device).

Each run's fixed dt is the initial CFL time step of fluidsim
(fluidsim/base/time_stepping/base.py:320-366: CFL / (max|u|/dx + max|v|/dy),
capped by deltat_max), evaluated on the device by the engine's tendency with
maxima and ``fsx_cfl_update``.
"""

import ctypes
import json
import os

import numpy as np

from . import _lib
from .operators.operators2d import OperatorsPseudoSpectral2D

_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "fluidsim_amd",
                              "ensemble_tuning.json")


def _operator(n, L, precision, coef_dealiasing, truncation_shape):
    params = OperatorsPseudoSpectral2D._create_default_params()
    params.oper.nx = params.oper.ny = n
    params.oper.Lx = params.oper.Ly = L
    params.oper.precision = precision
    params.oper.coef_dealiasing = coef_dealiasing
    params.oper.truncation_shape = truncation_shape
    return OperatorsPseudoSpectral2D(params)


class _Group:
    """Runs of one size: batch plan and state arrays (nb, n, n/2 + 1)."""

    def __init__(self, n, runs, ens):
        import torch

        self.n = n
        self.runs = runs  # indices into the ensemble's run list
        self.nb = len(runs)
        self.oper = _operator(n, ens.L, ens.precision, ens.coef_dealiasing,
                              ens.truncation_shape)
        oper = self.oper
        mask = oper.where_dealiased if oper._has_to_dealiase else None
        self._mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        h = ctypes.c_void_p()
        prec = oper.precision
        _lib.call("fsx_plan_create_batch", ctypes.byref(h), n, n, float(ens.L), float(ens.L),
                  prec, None if self._mask is None else self._mask.ctypes.data_as(
                      ctypes.c_void_p), self.nb)
        self.handle = h
        shape = (self.nb,) + tuple(oper.shapeK)
        ct = oper.dtype_complex
        self.S = torch.zeros(shape, dtype=ct, device="cuda")
        self.tmp = torch.zeros_like(self.S)
        self.t1 = torch.zeros_like(self.S)
        self.lines = 0
        self.row = 0
        self.stream = None  # the branch stream of the group (set by the ensemble)

    def close(self):
        if self.handle:
            _lib.call("fsx_plan_destroy", self.handle)
            self.handle = None


class EnsembleNS2D:
    """Independent ns2d runs of mixed sizes, advanced together.

    ``runs``: sequence of dicts with ``n`` (grid points per side, power of
    two), ``rot_fft`` (initial vorticity, half spectrum (n, n/2 + 1)),
    ``nu_2`` and optionally ``nu_4``, ``nu_8``, ``nu_m4`` and ``deltat``
    (default: the initial CFL time step with ``cfl_coef``).
    """

    def __init__(self, runs, precision="float32", L=2 * np.pi, coef_dealiasing=2.0 / 3,
                 truncation_shape="cubic", beta=0.0, scheme="RK4", cfl_coef=0.5,
                 deltat_max=0.2, autotune=True, tune_cache=None, stream=None):
        import torch

        if scheme not in ("RK4", "RK2"):
            raise ValueError('scheme must be "RK4" or "RK2"')
        self.precision = precision
        self.L = float(L)
        self.coef_dealiasing = coef_dealiasing
        self.truncation_shape = truncation_shape
        self.beta = float(beta)
        self.nstages = 4 if scheme == "RK4" else 2
        self.runs = [dict(r) for r in runs]
        if not self.runs:
            raise ValueError("empty ensemble")
        # a created stream: HIP graphs cannot capture the null stream
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self._sp = _lib.stream_ptr(self.stream)
        sizes = sorted({int(r["n"]) for r in self.runs})
        self.groups = []
        self.where = [None] * len(self.runs)
        for n in sizes:
            idx = [i for i, r in enumerate(self.runs) if int(r["n"]) == n]
            g = _Group(n, idx, self)
            for pos, i in enumerate(idx):
                self.where[i] = (g, pos)
            self.groups.append(g)
        # one branch stream per group beyond the first (default priority)
        self.groups[0].stream = self.stream
        for g in self.groups[1:]:
            g.stream = torch.cuda.Stream()
        # the groups' buffers were zero-filled on the current stream
        self.stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self.stream):
            for g in self.groups:
                self._load_group(g, cfl_coef, deltat_max)
        self.stream.synchronize()
        self.it = 0
        self._graph = None
        if autotune:
            self.autotune(tune_cache)

    # ------------------------------------------------------------ set-up
    def _load_group(self, g, cfl_coef, deltat_max):
        import torch

        oper = g.oper
        dts, visc = [], []
        work = torch.zeros(oper.shapeK, dtype=oper.dtype_complex, device="cuda")
        for pos, i in enumerate(g.runs):
            r = self.runs[i]
            rot = np.asarray(r["rot_fft"])
            if rot.shape != tuple(oper.shapeK):
                raise ValueError(f"run {i}: rot_fft of shape {rot.shape}, expected {oper.shapeK}")
            g.S[pos].copy_(torch.from_numpy(np.ascontiguousarray(rot)).to(g.S.dtype))
            oper.dealiasing(g.S[pos])
            dt = r.get("deltat")
            if dt is None:
                dt = self._cfl_deltat(oper, g.S[pos], work, cfl_coef, deltat_max)
            r["deltat"] = float(dt)
            dts.append(float(dt))
            visc += [float(r.get("nu_2", 0.0)), float(r.get("nu_4", 0.0)),
                     float(r.get("nu_8", 0.0)), float(r.get("nu_m4", 0.0))]
        _lib.call("fsx_batch_set_params", g.handle, _lib.doubles(dts), _lib.doubles(visc))

    def _cfl_deltat(self, oper, rot, work, cfl_coef, deltat_max):
        h = oper.plan.handle
        sp = self._sp
        _lib.call("fsx_set_dt", h, 0.0, sp)
        _lib.call("fsx_reset_maxima", h, sp)
        _lib.call("fsx_ns2d_tendency", h, _lib.dptr(rot), _lib.dptr(work), self.beta, 1, sp)
        _lib.call("fsx_cfl_update", h, 2, float(cfl_coef), float(deltat_max), sp)
        out = (ctypes.c_double * 6)()
        _lib.call("fsx_get_scalars", h, out, sp)
        return out[4]

    # ------------------------------------------------------------ steps
    def _issue_group(self, g, sp):
        for stage in range(self.nstages):
            _lib.call("fsx_ns2d_batch_rk_stage", g.handle, self.nstages, stage, _lib.dptr(g.S),
                      _lib.dptr(g.tmp), _lib.dptr(g.t1), self.beta, sp)

    def _issue_step(self, serial=False):
        """One step of every group: the first group on the ensemble's
        stream, the others on their branch streams, forked from it and
        joined back to it by events (under capture: parallel graph
        branches); serial: every group on the ensemble's stream."""
        import torch

        if serial:
            for g in self.groups:
                self._issue_group(g, self._sp)
            return
        main = self.stream
        fork = torch.cuda.Event()
        fork.record(main)
        for g in self.groups[1:]:
            g.stream.wait_event(fork)
            self._issue_group(g, _lib.stream_ptr(g.stream))
        self._issue_group(self.groups[0], self._sp)
        for g in self.groups[1:]:
            join = torch.cuda.Event()
            join.record(g.stream)
            main.wait_event(join)

    def _capture(self):
        import torch

        self.stream.synchronize()
        h = ctypes.c_void_p()
        _lib.call("fsx_graph_begin", self._sp)
        try:
            self._issue_step()
        finally:
            rc = _lib.get_lib().fsx_graph_end(self._sp, ctypes.byref(h))
        _lib.check(rc, "fsx_graph_end")
        self._graph = h
        torch.cuda.synchronize()

    def step(self, nsteps=1, graph=True, serial=False):
        """Advance every run by nsteps steps (one graph launch per step);
        without the graph, serial issues the groups one after the other on
        one stream (per-launch timing)."""
        if nsteps <= 0:
            return
        if graph and self._graph is None:
            # one eager step first: first launches set kernel attributes
            self._issue_step()
            nsteps -= 1
            self.it += 1
            self._capture()
        for _ in range(nsteps):
            if graph:
                _lib.call("fsx_graph_launch", self._graph, self._sp)
            else:
                self._issue_step(serial=serial)
            self.it += 1

    def synchronize(self):
        self.stream.synchronize()

    def check_nan(self):
        """ValueError if a NaN appeared in any run (pseudo_spect.py:241-243)."""
        self.synchronize()
        for g in self.groups:
            out = (ctypes.c_double * 6)()
            _lib.call("fsx_get_scalars", g.handle, out, self._sp)
            if out[5] != 0:
                raise ValueError(f"nan at it = {self.it} in the {g.n}^2 group")

    # ------------------------------------------------------------ access
    def get_state(self, i):
        """Vorticity (half spectrum) of run i as a numpy array."""
        g, pos = self.where[i]
        self.synchronize()
        return g.S[pos].cpu().numpy()

    def t(self, i):
        return self.it * self.runs[i]["deltat"]

    # ------------------------------------------------------------ tuning
    def _time_group(self, g, reps=3):
        import torch

        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(self.stream):
            S0 = g.S.clone()
            self._issue_group(g, self._sp)  # warm
            e0.record(self.stream)
            for _ in range(reps):
                self._issue_group(g, self._sp)
            e1.record(self.stream)
            g.S.copy_(S0)
        self.stream.synchronize()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # row passes the autotuner tries (fsx_batch_set_row): register rows with
    # radix-8 stages, with radix-4 stages (FP32 only), the LDS row pass
    ROW_MODES = {0: "register radix-8", 1: "register radix-4", 2: "LDS radix 8/4/2"}

    def autotune(self, cache=None):
        """Pick, per group, the lines per tile of the y-passes and the row
        pass by timing the group's RK step for every pair; choices cached in
        JSON keyed by size, batch, precision, scheme and device name."""
        import torch

        path = cache or os.environ.get("FLUIDSIM_AMD_TUNE_CACHE", _DEFAULT_CACHE)
        try:
            with open(path) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
        dev = torch.cuda.get_device_name(0)
        changed = False
        self.tuning = {}
        rows = [0, 1, 2] if self.precision == "float32" else [0, 2]
        for g in self.groups:
            key = f"{dev}|ns2d-v2|{g.n}|{g.nb}|{self.precision}|{self.nstages}"
            if key in table:
                lines, row = int(table[key]["lines"]), int(table[key]["row"])
            else:
                best, lines, row, times = None, 0, 0, {}
                cands = [0] + [1 << k for k in range(0, 8) if (1 << k) <= g.n]
                for r in rows:
                    _lib.call("fsx_batch_set_row", g.handle, r)
                    for cand in cands:
                        _lib.call("fsx_batch_set_lines", g.handle, cand)
                        ms = self._time_group(g)
                        times[f"{cand}/{r}"] = ms
                        if best is None or ms < best:
                            best, lines, row = ms, cand, r
                table[key] = {"lines": lines, "row": row, "row_pass": self.ROW_MODES[row],
                              "ms_per_step": best, "candidates (lines/row)": times}
                changed = True
            _lib.call("fsx_batch_set_lines", g.handle, lines)
            _lib.call("fsx_batch_set_row", g.handle, row)
            g.lines, g.row = lines, row
            self.tuning[g.n] = table[key]
        if changed:
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                with open(path, "w") as f:
                    json.dump(table, f, indent=1)
            except OSError:
                pass

    def close(self):
        if self._graph is not None:
            _lib.call("fsx_graph_destroy", self._graph)
            self._graph = None
        for g in self.groups:
            g.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
