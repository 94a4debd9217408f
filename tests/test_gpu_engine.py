"""GPU parity of the HIP engine against the FP64 oracle, through the C-ABI.

Tolerances: FP64 results agree with the oracle to round-off (relative L2
<= 1e-13 for single transforms / tendencies); FP32 to 2e-6 relative L2.
"""

import ctypes

import numpy as np
import pytest

from fluidsim_amd import _lib
from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3D
from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft

from .gpu_helpers import (
    to_dev_complex, to_dev_real, zeros_complex, zeros_real, host, make_plan,
    rel_l2,
)

pytestmark = pytest.mark.gpu

TOL = {8: 1e-13, 4: 2e-6}


def S():
    return _lib.stream_ptr()


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("shape", [(16, 16), (32, 64), (256, 256), (8, 512), (2048, 16),
                                   (8, 1024), (4, 2048), (1024, 8)])
def test_fft_ifft_2d(prec, shape):
    ny, nx = shape
    print(f"case 2D fft prec={prec} shape={shape}", flush=True)
    oper = Oper2D(nx, ny, 2 * np.pi, 3.0)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2,) + oper.shapeX)
    plan = make_plan(oper, prec, 1, dealias=False)
    xd = to_dev_real(x, prec)
    kd = zeros_complex((2,) + oper.shapeK, prec)
    _lib.call("fsx_fft", plan.handle, _lib.dptr(xd), _lib.dptr(kd), 2, S())
    ref = np.array([oper.fft(a) for a in x])
    assert rel_l2(host(kd), ref) < TOL[prec]
    # inverse of a non-Hermitian spectrum (numpy c2r semantics)
    k = ref + 0.01 * (rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape))
    kd2 = to_dev_complex(k, prec)
    xo = zeros_real((2,) + oper.shapeX, prec)
    _lib.call("fsx_ifft", plan.handle, _lib.dptr(kd2), _lib.dptr(xo), 2, S())
    refx = np.array([oper.ifft(a) for a in k])
    assert rel_l2(host(xo), refx) < TOL[prec]
    # input untouched
    assert np.array_equal(host(kd2), k.astype(host(kd2).dtype))


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 8, 32), (64, 64, 64), (4, 32, 128),
                                   (4, 8, 1024), (1024, 4, 8), (8, 1024, 4)])
def test_fft_ifft_3d(prec, shape):
    nz, ny, nx = shape
    print(f"case 3D fft prec={prec} shape={shape}", flush=True)
    oper = Oper3D(nx, ny, nz, 2 * np.pi, 3.0, 5.0)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(oper.shapeX)
    plan = make_plan(oper, prec, 1, dealias=False)
    xd = to_dev_real(x, prec)
    kd = zeros_complex(oper.shapeK, prec)
    _lib.call("fsx_fft", plan.handle, _lib.dptr(xd), _lib.dptr(kd), 1, S())
    ref = oper.fft(x)
    assert rel_l2(host(kd), ref) < TOL[prec]
    k = ref + 0.01 * (rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape))
    kd2 = to_dev_complex(k, prec)
    xo = zeros_real(oper.shapeX, prec)
    _lib.call("fsx_ifft", plan.handle, _lib.dptr(kd2), _lib.dptr(xo), 1, S())
    assert rel_l2(host(xo), oper.ifft(k)) < TOL[prec]


def _ns2d_case(n, seed=0):
    oper = Oper2D(n, n)
    sim = NS2D(oper, nu_2=1e-3)
    sim.init_from_rotfft(ns2d_noise_rot_fft(n, n, seed=seed, kmax=min(16, n / 4)))
    return oper, sim


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("n", [32, 256, 1024, 2048])
def test_ns2d_tendency(prec, n):
    print(f"case ns2d tendency prec={prec} n={n}", flush=True)
    oper, sim = _ns2d_case(n)
    ref = sim.tendencies_nonlin(sim.state_spect)[0]
    plan = make_plan(oper, prec, 4)
    rot = to_dev_complex(sim.state_spect[0], prec)
    out = zeros_complex(oper.shapeK, prec)
    _lib.call("fsx_reset_maxima", plan.handle, S())
    _lib.call("fsx_ns2d_tendency", plan.handle, _lib.dptr(rot), _lib.dptr(out),
              0.0, 1, S())
    assert rel_l2(host(out), ref) < (1e-12 if prec == 8 else 2e-5)
    sc = (ctypes.c_double * 6)()
    _lib.call("fsx_get_scalars", plan.handle, sc, S())
    ph = sim.state_phys()
    assert sc[0] == pytest.approx(np.abs(ph["ux"]).max(), rel=1e-12 if prec == 8 else 1e-5)
    assert sc[1] == pytest.approx(np.abs(ph["uy"]).max(), rel=1e-12 if prec == 8 else 1e-5)


@pytest.mark.parametrize("n", [256, 2048])
def test_ns2d_tile_settings_bit_identical(n):
    """fsx_plan_set_tiles changes only the tiling of the y-passes: the
    inverse with one output field per workgroup or all four, any lines per
    tile, give the same tendency bit for bit (FP32)."""
    print(f"case ns2d tile settings n={n}", flush=True)
    oper, sim = _ns2d_case(n)
    plan = make_plan(oper, 4, 4)
    rot = to_dev_complex(sim.state_spect[0], 4)
    outs = []
    for split, inv, fwd in ((0, 0, 0), (1, 0, 0), (1, 2, 1), (0, 8, 4)):
        _lib.call("fsx_plan_set_tiles", plan.handle, inv, fwd, split)
        out = zeros_complex(oper.shapeK, 4)
        _lib.call("fsx_ns2d_tendency", plan.handle, _lib.dptr(rot), _lib.dptr(out), 0.0, 0, S())
        outs.append(host(out))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def _ns3d_case(shape, seed=0):
    nz, ny, nx = shape
    oper = Oper3D(nx, ny, nz)
    sim = NS3D(oper, nu_2=1e-3)
    vv = ns3d_spectrum_vv_fft(nx, ny, nz, seed=seed)
    sim.init_from_vxvyvzfft(*vv)
    return oper, sim


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 16, 32), (64, 64, 64), (4, 8, 1024),
                                   (1024, 8, 4), (8, 1024, 8)])
def test_ns3d_tendency(prec, shape):
    print(f"case ns3d tendency prec={prec} shape={shape}", flush=True)
    oper, sim = _ns3d_case(shape)
    ref = sim.tendencies_nonlin(sim.state_spect)
    plan = make_plan(oper, prec, 6)
    v = to_dev_complex(sim.state_spect, prec)
    out = zeros_complex((3,) + oper.shapeK, prec)
    _lib.call("fsx_reset_maxima", plan.handle, S())
    _lib.call("fsx_ns3d_tendency", plan.handle, _lib.dptr(v), _lib.dptr(out),
              0, 0.0, 1, S())
    _lib.call("fsx_project", plan.handle, _lib.dptr(out[0]), _lib.dptr(out[1]),
              _lib.dptr(out[2]), 1, S())
    assert rel_l2(host(out), ref) < (1e-12 if prec == 8 else 2e-5)
    sc = (ctypes.c_double * 6)()
    _lib.call("fsx_get_scalars", plan.handle, sc, S())
    ph = sim.state_phys()
    for i, k in enumerate(("vx", "vy", "vz")):
        assert sc[i] == pytest.approx(np.abs(ph[k]).max(), rel=1e-12 if prec == 8 else 1e-5)


def _run_engine_rk4(plan, S_host, prec, nvar, dts, visc, tendency, project=0):
    S_d = to_dev_complex(S_host, prec)
    tmp = zeros_complex(S_d.shape, prec)
    t1 = zeros_complex(S_d.shape, prec)
    N = zeros_complex(S_d.shape, prec)
    vc = _lib.doubles(visc)
    for dt in dts:
        _lib.call("fsx_set_dt", plan.handle, float(dt), S())
        src = S_d
        for stage in range(4):
            tendency(src, N)
            _lib.call("fsx_rk_stage", plan.handle, 4, stage, nvar, _lib.dptr(S_d),
                      _lib.dptr(tmp), _lib.dptr(t1), _lib.dptr(N), project, 0, vc, S())
            src = t1
    return host(S_d)


@pytest.mark.parametrize("prec", [8, 4])
def test_ns2d_rk4_steps(prec):
    print(f"case ns2d rk4 prec={prec}", flush=True)
    oper, sim = _ns2d_case(64, seed=4)
    S0 = sim.state_spect.copy()
    dts = [0.01, 0.01, 0.012]
    sim.run_with_dts(dts)
    plan = make_plan(oper, prec, 4)

    def tend(src, N):
        _lib.call("fsx_ns2d_tendency", plan.handle, _lib.dptr(src), _lib.dptr(N),
                  0.0, 0, S())

    out = _run_engine_rk4(plan, S0, prec, 1, dts, [1e-3, 0, 0, 0], tend)
    assert rel_l2(out, sim.state_spect) < (1e-12 if prec == 8 else 1e-5)


@pytest.mark.parametrize("prec", [8, 4])
def test_ns3d_rk4_steps(prec):
    print(f"case ns3d rk4 prec={prec}", flush=True)
    oper, sim = _ns3d_case((32, 32, 32), seed=5)
    S0 = sim.state_spect.copy()
    dts = [0.01, 0.01]
    sim.run_with_dts(dts)
    plan = make_plan(oper, prec, 6)

    def tend(src, N):
        _lib.call("fsx_ns3d_tendency", plan.handle, _lib.dptr(src), _lib.dptr(N),
                  0, 0.0, 0, S())

    out = _run_engine_rk4(plan, S0, prec, 3, dts, [1e-3, 0, 0, 0], tend, project=1)
    assert rel_l2(out, sim.state_spect) < (1e-12 if prec == 8 else 1e-5)


@pytest.mark.parametrize("prec", [8, 4])
def test_sums_and_spectrum(prec):
    oper, sim = _ns2d_case(64, seed=1)
    plan = make_plan(oper, prec, 4)
    rot = to_dev_complex(sim.state_spect[0], prec)
    out = (ctypes.c_double * 6)()
    visc = _lib.doubles([1e-3, 0, 0, 0])
    _lib.call("fsx_sums", plan.handle, 0, _lib.dptr(rot), None, None, None, 0.0,
              visc, out, S())
    m = sim.spatial_means()
    tol = 1e-12 if prec == 8 else 1e-6
    assert out[0] == pytest.approx(m["E"], rel=tol)
    assert out[1] == pytest.approx(m["Z"], rel=tol)
    assert out[2] == pytest.approx(m["epsK"], rel=tol)
    assert out[4] == pytest.approx(m["epsZ"], rel=tol)


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_projections_vs_oracle(kind):
    oper = Oper3D(16, 8, 32, 2 * np.pi, 3.0, 5.0)
    rng = np.random.default_rng(kind)
    v = np.array([oper.fft(rng.standard_normal(oper.shapeX)) for _ in range(3)])
    ref = v.copy()
    {1: oper.project_perpk3d, 2: oper.project_toroidal,
     3: oper.project_poloidal}[kind](ref[0], ref[1], ref[2])
    plan = make_plan(oper, 8, 6)
    d = to_dev_complex(v, 8)
    _lib.call("fsx_project", plan.handle, _lib.dptr(d[0]), _lib.dptr(d[1]),
              _lib.dptr(d[2]), kind, S())
    assert rel_l2(host(d), ref) < 1e-14


@pytest.mark.parametrize("kind", [1, 2])
def test_ns3d_rk4_projection_kinds(kind):
    proj = {1: None, 2: "toroidal"}[kind]
    oper = Oper3D(16, 16, 16)
    sim = NS3D(oper, nu_2=1e-3, projection=proj)
    vv = ns3d_spectrum_vv_fft(16, seed=3)
    sim.state_spect[:] = np.array(vv)
    sim.project_state_spect(sim.state_spect)
    S0 = sim.state_spect.copy()
    ref_tend = sim.tendencies_nonlin(S0.copy())
    plan = make_plan(oper, 8, 6)
    # tendency + projection
    v = to_dev_complex(S0, 8)
    out = zeros_complex((3,) + oper.shapeK, 8)
    _lib.call("fsx_ns3d_tendency", plan.handle, _lib.dptr(v), _lib.dptr(out), 0, 0.0, 0, S())
    _lib.call("fsx_project", plan.handle, _lib.dptr(out[0]), _lib.dptr(out[1]),
              _lib.dptr(out[2]), kind, S())
    assert rel_l2(host(out), ref_tend) < 1e-12
    sim.run_with_dts([0.01])

    def tend(src, N):
        _lib.call("fsx_ns3d_tendency", plan.handle, _lib.dptr(src), _lib.dptr(N),
                  0, 0.0, 0, S())

    got = _run_engine_rk4(plan, S0, 8, 3, [0.01], [1e-3, 0, 0, 0], tend, project=kind)
    assert rel_l2(got, sim.state_spect) < 1e-12
