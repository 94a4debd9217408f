"""CPU tests of the state-file container (fluidsim_amd.base.phys_fields):
the tree of fluidsim/util/output.py:47-123 (state_phys/<key>, attrs time and
it, info_simul/params) written and read back without a GPU, the file-name
rule of fluidsim/base/output/phys_fields.py:137-191, and the parameter tree
rebuilt over the defaults."""

import os
import types

import numpy as np
import pytest

from fluidsim_amd.base import phys_fields as PF
from fluidsim_amd.params import Parameters


def _fake_sim(tmp_path):
    params = Parameters(tag="params")
    params._set_attribs({"nu_2": 1e-3, "NEW_DIR_RESULTS": True})
    params._set_child("oper", attribs={"nx": 12, "ny": 9, "Lx": 6.0, "Ly": 4.5})
    params._set_child("time_stepping", attribs={"USE_T_END": True, "t_end": 25.0})
    params._set_child("output")
    params.output._set_child("phys_fields", attribs={"file_with_it": False})

    class Simul:
        class InfoSolver:
            _extenders = []

    Simul.__module__ = "fluidsim_amd.solvers.ns2d.solver"
    sim = Simul()
    sim.params = params
    sim.oper = types.SimpleNamespace(axes=("y", "x"))
    sim.info_solver = types.SimpleNamespace(short_name="NS2D")
    sim.output = types.SimpleNamespace(name_run="NS2D_test", path_run=str(tmp_path))
    return sim


def test_save_and_read_back(tmp_path):
    sim = _fake_sim(tmp_path)
    rng = np.random.default_rng(0)
    arrays = {"rot": rng.standard_normal((9, 12)), "ux": rng.standard_normal((9, 12))}
    name = PF.name_file(3.25, 17, sim.params, str(tmp_path), 0.5)
    # width from t_end = 25: int(log10 25) + 7 = 8 characters
    assert name == "state_phys_t0003.250.npz"
    path = tmp_path / name
    PF.save_file(path, arrays, "state_phys", sim, 3.25, 17)
    f = PF.StateFile(path)
    assert sorted(f.keys) == ["rot", "ux"]
    np.testing.assert_array_equal(f["rot"], arrays["rot"])
    assert f.state_attrs["time"] == 3.25 and f.state_attrs["it"] == 17
    assert f.state_attrs["what"] == "obj state_phys for fluidsim"
    assert f.attrs["axes"] == ["y", "x"] and f.attrs["name_solver"] == "NS2D"
    assert f.params["oper"]["nx"] == 12 and f.params["SAVE"] == 1
    assert f.solver["module_name"] == "fluidsim_amd.solvers.ns2d.solver"
    f.close()
    # a second file: the last one and the one closest to a time
    PF.save_file(tmp_path / PF.name_file(7.5, 40, sim.params, str(tmp_path), 0.5),
                 arrays, "state_phys", sim, 7.5, 40)
    assert PF.path_file_from_time_approx(tmp_path).endswith("t0007.500.npz")
    assert PF.path_file_from_time_approx(tmp_path, 3.0).endswith("t0003.250.npz")
    # short periods name the iteration (phys_fields.py:167-170)
    assert PF.name_file(1.0, 5, sim.params, str(tmp_path / "x"), 0.0005).endswith(
        "_it=5.npz")


def test_not_a_state_file(tmp_path):
    p = tmp_path / "junk.npz"
    np.savez(p, a=np.zeros(3))
    with pytest.raises(ValueError, match="params"):
        PF.StateFile(p)
    q = tmp_path / "junk.txt"
    q.write_text("not a zip")
    with pytest.raises(ValueError, match="really a state file"):
        PF.StateFile(q)


def test_params_tree_over_defaults():
    d = Parameters(tag="params")
    d._set_attribs({"nu_2": 0.0, "nu_8": 0.0})
    d._set_child("oper", attribs={"nx": 48, "ny": 48})
    tree = {"nu_2": 1e-3, "new_key": 2, "oper": {"nx": 12}, "forcing": {"enable": True}}
    d._update_from_dict_tree(tree)
    assert d.nu_2 == 1e-3 and d.nu_8 == 0.0 and d.new_key == 2
    assert d.oper.nx == 12 and d.oper.ny == 48 and d.forcing.enable is True
    with pytest.raises(ValueError):
        d._update_from_dict_tree({"oper": 3})
    assert os.path.basename(PF.__file__) == "phys_fields.py"
