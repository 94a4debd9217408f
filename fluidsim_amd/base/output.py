"""Output (fluidsim/base/output/base.py:60-521, subset).

Periodic printing (print_stdout, periods_print.print_stdout) and the
specific outputs of fluidsim_amd.base.diagnostics, computed on the device:
spatial_means, spectra and spect_energy_budg (periods_save.*), created at
post_init, first saved when the state is initialised, then every period.
Files (stdout.txt, params.json, spatial_means.txt, spectra / budget .npz) go
to path_run when params.output.HAS_TO_SAVE, and so do the physical-field
state files of fluidsim_amd.base.phys_fields (periods_save.phys_fields, and
one at the end of a run that saves).  Plots are not implemented.
"""

import json
import os
from datetime import datetime

import numpy as np

from .. import _lib


class OutputBasePseudoSpectral:
    @staticmethod
    def _complete_params_with_default(params, info_solver=None):
        params._set_child(
            "output",
            attribs={
                "HAS_TO_SAVE": True,
                "sub_directory": "",
                "ONLINE_PLOT_OK": False,
            },
        )
        params.output._set_child(
            "periods_save",
            attribs={"phys_fields": 0, "spatial_means": 0, "spectra": 0,
                     "spect_energy_budg": 0},
        )
        for tag in ("spatial_means", "spectra", "spect_energy_budg"):
            params.output._set_child(tag, attribs={"HAS_TO_PLOT_SAVED": False})
        params.output._set_child("periods_print", attribs={"print_stdout": 1.0})
        params.output._set_child("periods_plot", attribs={"phys_fields": 0})
        params.output._set_child("phys_fields",
                                 attribs={"field_to_plot": "", "file_with_it": False})

    def __init__(self, sim):
        self.sim = sim
        self.params = sim.params
        self.oper = sim.oper
        p = self.params
        name = "_".join(
            s for s in (
                sim.info_solver.short_name,
                p.short_name_type_run,
                self.oper.produce_str_describing_oper(),
                datetime.now().strftime("%Y-%m-%d_%H-%M-%S"),
            ) if s
        )
        self.name_run = name
        self.name_solver = sim.info_solver.short_name
        self._has_to_save = bool(p.output.HAS_TO_SAVE)
        base = os.environ.get("FLUIDSIM_PATH", os.path.expanduser("~/Sim_data"))
        if not p.NEW_DIR_RESULTS and getattr(p, "path_run", ""):
            # restart in the directory of the run (fluidsim/util/util.py:384-385)
            self.path_run = str(p.path_run)
            self.name_run = os.path.basename(os.path.normpath(self.path_run))
        else:
            self.path_run = os.path.join(base, p.output.sub_directory, name)
            if self._has_to_save:
                # a new directory per run: runs started within the same
                # second get a numbered suffix
                i = 0
                while True:
                    try:
                        os.makedirs(self.path_run)
                        break
                    except FileExistsError:
                        i += 1
                        self.name_run = f"{name}_{i}"
                        self.path_run = os.path.join(base, p.output.sub_directory,
                                                     self.name_run)
        if "path_run" in p:
            p.path_run = self.path_run
        else:
            p._set_attrib("path_run", self.path_run)
        self._t_last = {}
        self._end_of_simul = False
        self._specific = []

    # ----------------------------------------------------------------
    # per-solver classes of the specific outputs (None: not available)
    _cls_spatial_means = None
    _cls_spectra = None
    _cls_spect_energy_budg = None

    def post_init(self):
        from .phys_fields import PhysFields

        for name in ("spatial_means", "spectra", "spect_energy_budg"):
            cls = getattr(self, "_cls_" + name)
            if cls is not None:
                obj = cls(self)
                setattr(self, name, obj)
                self._specific.append(obj)
        self.phys_fields = PhysFields(self)
        self._specific.append(self.phys_fields)
        if self._has_to_save:
            os.makedirs(self.path_run, exist_ok=True)
            with open(os.path.join(self.path_run, "params.json"), "w") as f:
                json.dump(self.params._make_dict_tree(), f, indent=1, default=str)
        self.print_stdout(
            f"*************************************\nfluidsim_amd solver "
            f"{self.sim.info_solver.short_name}\nsim.name_run: {self.name_run}"
        )

    def print_stdout(self, to_print, end="\n"):
        print(to_print, end=end, flush=True)
        if self._has_to_save:
            with open(os.path.join(self.path_run, "stdout.txt"), "a") as f:
                f.write(to_print + end)

    def init_with_initialized_state(self):
        """base.py:263-300: first CFL, then outputs at t = 0."""
        if getattr(self, "_has_been_initialized_with_state", False):
            return
        self._has_been_initialized_with_state = True
        if self.params.time_stepping.USE_CFL:
            self.sim.time_stepping.compute_time_increment_CLF()
        self._do_outputs(force=True)
        for obj in self._specific:
            obj._init_files()

    def outputs_due(self):
        """Whether a print or a periodic save is due at the current time (such
        outputs read dt and the forcing of the coming step)."""
        t = self.sim.time_stepping.t
        per = self.params.output.periods_print.print_stdout
        last = self._t_last.get("print")
        if per and (last is None or t - last >= per):
            return True
        for obj in self._specific:
            if obj.due(t):
                return True
        return False

    def one_time_step(self):
        self._do_outputs(force=False)
        for obj in self._specific:
            obj._online_save()

    def _due(self, key, period, t, force):
        if not period:
            return False
        last = self._t_last.get(key)
        if force or last is None or t - last >= period:
            self._t_last[key] = t
            return True
        return False

    def _do_outputs(self, force):
        ts = self.sim.time_stepping
        t = ts.t
        p = self.params.output
        if self._due("print", p.periods_print.print_stdout, t, force):
            self.print_stdout(
                f"it = {ts.it:6d} ; t = {t:12.6g} ; deltat  = {ts.deltat:10.5g}"
            )

    def compute_spatial_means(self):
        """One spatial-means record (device sums)."""
        return self.spatial_means.compute()

    def load_spatial_means(self):
        return self.spatial_means.load()

    def end_of_simul(self, total_time=None):
        """fluidsim/base/output/base.py:354-361: a run that saves writes its
        final state file (unless one was written at this time)."""
        self._end_of_simul = True
        if self._has_to_save and self.phys_fields.t_last_save < self.sim.time_stepping.t:
            self.phys_fields.save()

    # ---------------------------------------------------------------- sums
    def _sums(self, kind, fields, N2=0.0):
        out = (_lib.ctypes.c_double * 6)()
        f = [(_lib.dptr(a) if a is not None else None) for a in fields]
        f += [None] * (4 - len(f))
        _lib.call("fsx_sums", self.oper.plan.handle, kind, f[0], f[1], f[2], f[3],
                  float(N2), _lib.doubles(self.sim.visc_coefs()), out,
                  _lib.stream_ptr())
        return list(out)
