"""GPU tests of the state files and restarts (fluidsim/base/output/
phys_fields.py:130-200, fluidsim/util/output.py:47-123,
fluidsim/base/init_fields.py:140-297, fluidsim/util/util.py:338-399).

A run that saves writes its final state file; load_for_restart gives the
parameters and Simul class of the run with init_fields.type = "from_file";
the resumed run continues from the file's state, time and iteration.  Bar:
an FP64 run stopped and resumed equals the uninterrupted run to round-off
(relative L2 <= 1e-13: the state passes through one inverse and one forward
transform), with the same time and iteration.
"""

import os

import numpy as np
import pytest

from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def fluidsim_path(tmp_path):
    old = os.environ.get("FLUIDSIM_PATH")
    os.environ["FLUIDSIM_PATH"] = str(tmp_path)
    yield tmp_path
    if old is None:
        del os.environ["FLUIDSIM_PATH"]
    else:
        os.environ["FLUIDSIM_PATH"] = old


def _ns3d(n, it_end, precision="float64"):
    from fluidsim_amd.solvers.ns3d.solver import Simul

    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.oper.precision = precision
    params.nu_2 = 1e-2
    params.time_stepping.USE_CFL = False
    params.time_stepping.deltat0 = 0.01
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = it_end
    params.init_fields.type = "in_script"
    params.output.periods_print.print_stdout = 0
    params.output.periods_save.spatial_means = 0.02
    sim = Simul(params)
    vv = ns3d_spectrum_vv_fft(n, seed=7)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    return sim


@pytest.mark.parametrize("n", [32, 24])
def test_ns3d_restart_equals_uninterrupted(fluidsim_path, n):
    from fluidsim_amd.base.phys_fields import StateFile, path_file_from_time_approx
    from fluidsim_amd.util import load_for_restart

    print(f"case ns3d restart n={n}", flush=True)
    full = _ns3d(n, 10)
    full.time_stepping.start()
    ref = full.state.state_spect.to_numpy()

    first = _ns3d(n, 5)
    first.time_stepping.start()
    path_run = first.output.path_run
    path = path_file_from_time_approx(path_run)
    f = StateFile(path)
    assert sorted(f.keys) == ["vx", "vy", "vz"]
    assert f.state_attrs["it"] == 5
    assert f.state_attrs["time"] == pytest.approx(0.05, rel=1e-12)
    assert f.attrs["axes"] == ["z", "y", "x"]
    assert f.params["oper"]["nx"] == n
    assert f["vx"].dtype == np.float64 and f["vx"].shape == (n, n, n)
    np.testing.assert_array_equal(f["vy"], first.state.state_phys.get_var("vy").cpu().numpy())
    f.close()

    params, Simul = load_for_restart(path_run)
    assert params.init_fields.type == "from_file"
    params.time_stepping.it_end = 10
    sim = Simul(params)
    assert sim.time_stepping.it == 5
    assert sim.output.path_run == path_run
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()
    err = rel_l2(got, ref)
    print(f"ns3d {n}^3 restart at it = 5: rel L2 vs uninterrupted = {err:.3e}")
    assert sim.time_stepping.it == 10
    assert sim.time_stepping.t == pytest.approx(full.time_stepping.t, rel=1e-13)
    assert err <= 1e-13
    # the resumed run wrote its final state file in the same directory
    assert path_file_from_time_approx(path_run).endswith("state_phys_t000.100.npz")


def test_ns2d_restart_and_grid_check(fluidsim_path):
    from fluidsim_amd.solvers.ns2d.solver import Simul
    from fluidsim_amd.util import load_for_restart

    n = 64

    def make(it_end):
        params = Simul.create_default_params()
        params.oper.nx = params.oper.ny = n
        params.nu_2 = 1e-3
        params.time_stepping.USE_CFL = False
        params.time_stepping.deltat0 = 0.01
        params.time_stepping.USE_T_END = False
        params.time_stepping.it_end = it_end
        params.init_fields.type = "in_script"
        params.output.periods_print.print_stdout = 0
        params.output.periods_save.phys_fields = 0.03
        sim = Simul(params)
        sim.state.init_statespect_from(rot_fft=ns2d_noise_rot_fft(n, n, seed=3, kmax=10))
        return sim

    full = make(12)
    full.time_stepping.start()
    first = make(6)
    first.time_stepping.start()
    files = sorted(os.listdir(first.output.path_run))
    # periodic saves (t = 0.03, 0.06 ... at the step outputs) and the final one
    assert sum(f.startswith("state_phys_t") for f in files) >= 2
    params, S2 = load_for_restart(first.output.path_run)
    params.time_stepping.it_end = 12
    sim = S2(params)
    sim.time_stepping.start()
    err = rel_l2(sim.state.get_var_numpy("rot_fft"), full.state.get_var_numpy("rot_fft"))
    print(f"ns2d {n}^2 restart: rel L2 = {err:.3e}")
    assert err <= 1e-13
    # a file of another grid is refused (init_fields.py:185-197)
    params, S2 = load_for_restart(first.output.path_run)
    params.oper.nx = 32
    with pytest.raises(ValueError, match="not a correct state"):
        S2(params)
