"""Physical-field state files: saving and loading for restarts
(fluidsim/base/output/phys_fields.py:130-200, fluidsim/util/output.py:47-123,
fluidsim/base/init_fields.py:140-297).

The reference writes ``state_phys_t{time}.nc`` / ``.h5`` with the HDF5 layout

    /                 attrs: "date saving", name_solver, name_run, axes
    /state_phys/<key> one dataset per physical state variable
    /state_phys       attrs: what, name_type_variables, time, it
    /info_simul/params the parameter tree

HDF5 libraries (h5py, h5netcdf, netCDF4) are not available to this
engine, so the same tree is stored in a numpy ``.npz`` container: member
``state_phys/<key>`` holds the array (plan precision, X layout), members
``.attrs``, ``state_phys/.attrs`` and ``info_simul/params`` hold JSON text
(0-d unicode arrays, readable with ``allow_pickle=False``), and
``info_simul/solver`` names the module and class of the Simul that wrote the
file.  The file name follows the reference's rule.
"""

import datetime
import glob
import json
import os

import numpy as np

EXT = "npz"


def _json_array(obj):
    return np.array(json.dumps(obj, default=str))


def save_file(path_file, state_phys_arrays, name_type_variables, sim, time, it,
              particular_attr=None):
    """Write one state file (fluidsim/util/output.py:47-123, single process).
    state_phys_arrays: dict key -> host ndarray (X layout)."""
    oper = sim.oper
    members = {}
    for k, a in state_phys_arrays.items():
        members["state_phys/" + k] = np.ascontiguousarray(a)
    members["state_phys/.attrs"] = _json_array({
        "what": "obj state_phys for fluidsim",
        "name_type_variables": name_type_variables,
        "time": float(time),
        "it": int(it),
    })
    root = {
        "date saving": str(datetime.datetime.now()),
        "name_solver": sim.info_solver.short_name,
        "name_run": sim.output.name_run,
        "axes": list(oper.axes),
    }
    if particular_attr is not None:
        root["particular_attr"] = particular_attr
    members[".attrs"] = _json_array(root)
    params = sim.params._make_dict_tree()
    params["SAVE"] = 1
    params["NEW_DIR_RESULTS"] = 1
    members["info_simul/params"] = _json_array(params)
    # the solver's own Simul class (an extended class derives from it) and
    # the extenders applied to it (turbulence models)
    base = next((c for c in type(sim).__mro__ if c.__module__.endswith(".solver")),
                type(sim))
    extenders = getattr(type(sim).InfoSolver, "_extenders", [])
    members["info_simul/solver"] = _json_array({
        "module_name": base.__module__, "class_name": base.__name__,
        "short_name": sim.info_solver.short_name,
        "extenders": [[c.__module__, c.__name__] for c in extenders],
    })
    tmp = str(path_file) + ".tmp.npz"
    np.savez(tmp, **members)
    os.replace(tmp, str(path_file))


class StateFile:
    """Read side of a state file (the tree of save_file)."""

    def __init__(self, path_file):
        self.path = str(path_file)
        try:
            self._npz = np.load(self.path, allow_pickle=False)
        except Exception as exc:
            raise ValueError(f"Is file {self.path} really a state file?") from exc
        names = set(self._npz.files)
        if "info_simul/params" not in names:
            raise ValueError(f"The file {self.path} does not contain a params object")
        if "state_phys/.attrs" not in names:
            raise ValueError(f"The file {self.path} does not contain a state_phys object")
        self.attrs = json.loads(str(self._npz[".attrs"]))
        self.state_attrs = json.loads(str(self._npz["state_phys/.attrs"]))
        self.params = json.loads(str(self._npz["info_simul/params"]))
        self.solver = (json.loads(str(self._npz["info_simul/solver"]))
                       if "info_simul/solver" in names else {})
        self.keys = [n[len("state_phys/"):] for n in self._npz.files
                     if n.startswith("state_phys/") and not n.endswith(".attrs")]

    def __getitem__(self, key):
        return self._npz["state_phys/" + key]

    def close(self):
        self._npz.close()


def name_file(time, it, params, path_run, period_save):
    """state_phys_t{time}.{ext}, width from t_end when USE_T_END
    (fluidsim/base/output/phys_fields.py:137-191)."""
    ts = params.time_stepping
    if ts.USE_T_END:
        existing = sorted(glob.glob(os.path.join(path_run, "state_phys*")))
        if existing:
            str_width = len(os.path.basename(existing[0])[12:-(len(EXT) + 1)])
        else:
            # digits of t_end, .3f, and 2 more for restarts
            str_width = int(np.log10(ts.t_end)) + 7
    else:
        str_width = 7
    str_it = ""
    if 0 < period_save < 0.001 or params.output.phys_fields.file_with_it:
        str_it = f"_it={it}"
    return f"state_phys_t{time:0{str_width}.3f}{str_it}.{EXT}"


class PhysFields:
    """Periodic saving of the physical state (PhysFieldsBase,
    fluidsim/base/output/phys_fields.py:36-200), without plotting."""

    _tag = "phys_fields"

    def __init__(self, output):
        self.output = output
        self.sim = output.sim
        self.params = output.params
        self.period_save = float(self.params.output.periods_save.phys_fields or 0.0)
        if self.period_save == 0:
            self.t_last_save = -np.inf
        else:
            self.t_last_save = self.sim.time_stepping.t

    def due(self, t):
        return False  # a state file does not read dt or the forcing

    def _init_files(self):
        pass

    def _has_to_online_save(self):
        # base.py:675-678
        return (self.sim.time_stepping.t + 1e-15) // self.period_save > \
            self.t_last_save // self.period_save

    def _online_save(self):
        if self.period_save and self._has_to_online_save():
            self.t_last_save = self.sim.time_stepping.t
            self.save()

    def save(self, particular_attr=None):
        """Write state_phys at the current time (skipped when a file of this
        time already holds the same iteration)."""
        sim = self.sim
        ts = sim.time_stepping
        path_run = self.output.path_run
        os.makedirs(path_run, exist_ok=True)
        name = name_file(ts.t, ts.it, self.params, path_run, self.period_save)
        path = os.path.join(path_run, name)
        if os.path.exists(path):
            f = StateFile(path)
            it_file = f.state_attrs.get("it")
            f.close()
            if it_file == ts.it:
                return
            name = f"state_phys_t{ts.t:07.3f}_it={ts.it}.{EXT}"
            path = os.path.join(path_run, name)
        self.output.print_stdout("save state_phys in file " + name)
        state = sim.state
        arrays = {k: state.state_phys.get_var(k).detach().cpu().numpy()
                  for k in state.keys_state_phys}
        save_file(path, arrays, "state_phys", sim, ts.t, ts.it, particular_attr)
        self.t_last_save = ts.t
        return path


def path_file_from_time_approx(path_dir, t_approx="last"):
    """State file of a run directory closest to t_approx
    (fluidsim/util/util.py:_path_file_from_time_approx)."""
    files = sorted(glob.glob(os.path.join(str(path_dir), "state_phys_t*." + EXT)))
    if not files:
        raise ValueError(f"No state file in {path_dir}")
    if t_approx is None or t_approx == "last":
        return files[-1]
    times = []
    for f in files:
        base = os.path.basename(f)[len("state_phys_t"):-(len(EXT) + 1)]
        times.append(float(base.split("_it=")[0]))
    return files[int(np.argmin(np.abs(np.array(times) - float(t_approx))))]
