"""Restart helpers (fluidsim/util/util.py:338-399)."""

import importlib
import os

from .base.phys_fields import StateFile, path_file_from_time_approx


def _path_dir(name_dir):
    if name_dir is None:
        return os.getcwd()
    name_dir = str(name_dir)
    if os.path.isdir(name_dir):
        return os.path.abspath(name_dir)
    base = os.environ.get("FLUIDSIM_PATH", os.path.expanduser("~/Sim_data"))
    cand = os.path.join(base, name_dir)
    if os.path.isdir(cand):
        return cand
    raise ValueError(f"No run directory {name_dir}")


def _simul_class(solver):
    """The Simul class recorded in a state file, extended with its extenders
    (fluidsim/util/util.py:_import_solver_from_path,
    _extend_simul_class_from_path)."""
    Simul = getattr(importlib.import_module(solver["module_name"]), solver["class_name"])
    extenders = [getattr(importlib.import_module(m), c)
                 for m, c in solver.get("extenders", [])]
    if extenders:
        from .extend_simul import extend_simul_class

        Simul = extend_simul_class(Simul, extenders)
    return Simul


def load_for_restart(name_dir=None, t_approx="last", merge_missing_params=False):
    """Load params and Simul for a restart.

    >>> params, Simul = load_for_restart(name_dir)
    >>> sim = Simul(params)

    The parameters are those saved in the state file (closest to t_approx,
    or the last one) over the solver's defaults (parameters the file lacks
    keep their default values, as with merge_missing_params); the run
    continues in its directory from the file's state, time and iteration.
    """
    path_dir = _path_dir(name_dir)
    path_file = path_file_from_time_approx(path_dir, t_approx)
    f = StateFile(path_file)
    try:
        if not f.solver:
            raise ValueError(f"The file {path_file} does not name its solver")
        Simul = _simul_class(f.solver)
        params = Simul.create_default_params()
        tree = dict(f.params)
        tree.pop("SAVE", None)
        params._update_from_dict_tree(tree)
    finally:
        f.close()
    params.path_run = path_dir
    params.NEW_DIR_RESULTS = False
    params.init_fields.type = "from_file"
    params.init_fields.from_file.path = path_file
    params.init_fields.modif_after_init = False
    return params, Simul
