"""GPU parity of the device resolution change (fsx_resize_fft) against the
oracle restatement of fill_field_fft_2d / fill_field_fft_3d, and of the
"from_simul" initialisation (fluidsim/base/init_fields.py:340-432).
Bit-exact: the kernel only copies values."""

import numpy as np
import pytest
import torch

from oracle import resize as R
from oracle.operators import Oper2D, Oper3D
from fluidsim_amd import _lib
from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft

pytestmark = pytest.mark.gpu


def dev_resize(a, shape_out, prec=8):
    src = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = torch.zeros((a.shape[0],) + tuple(shape_out), dtype=src.dtype, device="cuda")
    si, so = a.shape[1:], tuple(shape_out)
    if len(si) == 2:
        si, so = (1,) + si, (1,) + so
    _lib.call("fsx_resize_fft", _lib.dptr(src), si[0], si[1], si[2], _lib.dptr(out),
              so[0], so[1], so[2], a.shape[0], prec, _lib.stream_ptr())
    return out.cpu().numpy()


@pytest.mark.parametrize("shapes", [((64, 33), (16, 9)), ((16, 9), (64, 33)),
                                    ((32, 17), (64, 9))])
def test_resize_2d_bit_exact(shapes):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2,) + shapes[0]) + 1j * rng.standard_normal((2,) + shapes[0])
    got = dev_resize(a, shapes[1])
    for f in range(2):
        ref = np.zeros(shapes[1], complex)
        R.fill_field_fft_2d(a[f], ref)
        assert np.array_equal(got[f], ref)


@pytest.mark.parametrize("shapes", [((32, 32, 17), (16, 8, 9)), ((8, 16, 5), (32, 32, 17)),
                                    ((16, 32, 9), (32, 16, 5))])
def test_resize_3d_bit_exact(shapes):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((3,) + shapes[0]) + 1j * rng.standard_normal((3,) + shapes[0])
    got = dev_resize(a, shapes[1])
    for f in range(3):
        ref = np.zeros(shapes[1], complex)
        R.fill_field_fft_3d(a[f], ref)
        assert np.array_equal(got[f], ref)


def test_from_simul_ns2d_changes_resolution():
    from fluidsim_amd.solvers.ns2d.solver import Simul

    def make(n):
        params = Simul.create_default_params()
        params.oper.nx = params.oper.ny = n
        params.oper.Lx = params.oper.Ly = 2 * np.pi
        params.output.HAS_TO_SAVE = False
        params.init_fields.type = "in_script"
        return params

    sim_in = Simul(make(128))
    rot = ns2d_noise_rot_fft(128, 128, seed=4, kmax=10)
    sim_in.state.init_statespect_from(rot_fft=rot)
    sim_in.time_stepping.t = 1.25
    params = make(64)
    params.init_fields.type = "from_simul"
    sim = Simul(params)
    sim.init_fields.get_state_from_simul(sim_in)
    got = sim.state.get_var_numpy("rot_fft")
    ref = np.zeros(Oper2D(64, 64).shapeK, complex)
    R.fill_field_fft_2d(sim_in.state.get_var_numpy("rot_fft"), ref)
    assert np.array_equal(got, ref)
    assert sim.time_stepping.t == 1.25
    # a band-limited field keeps its physical values on the coarse points
    x = sim.state.state_phys.get_var("rot")
    x = x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x)
    x_in = Oper2D(128, 128).ifft(sim_in.state.get_var_numpy("rot_fft"))
    np.testing.assert_allclose(x, x_in[::2, ::2], atol=1e-12)


def test_from_simul_ns3d_fp32():
    from fluidsim_amd.solvers.ns3d.solver import Simul

    def make(n):
        params = Simul.create_default_params()
        params.oper.nx = params.oper.ny = params.oper.nz = n
        params.oper.precision = "float32"
        params.output.HAS_TO_SAVE = False
        params.init_fields.type = "in_script"
        return params

    sim_in = Simul(make(32))
    vv = ns3d_spectrum_vv_fft(32, seed=5, dtype=np.complex64)
    sim_in.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    params = make(64)
    params.init_fields.type = "from_simul"
    sim = Simul(params)
    sim.init_fields.get_state_from_simul(sim_in)
    got = sim.state.state_spect.to_numpy()
    src = sim_in.state.state_spect.to_numpy()
    for f in range(3):
        ref = np.zeros(Oper3D(64, 64, 64).shapeK, np.complex64)
        R.fill_field_fft_3d(src[f], ref)
        assert np.array_equal(got[f], ref)
