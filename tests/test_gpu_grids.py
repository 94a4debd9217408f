"""GPU parity on grids that are not powers of two (mixed-radix passes).

fluidsim accepts any FFTW size.  The reference's default grid is
nx = ny (= nz) = 48 (fluidsim/operators/operators2d.py:105-106,
fluidsim/operators/operators3d.py:159-161), its operator tests use nh = 9 and
12 (fluidsim/operators/test/test_operators2d.py:21-30, 121) and its ns3d
tests 16 x 12 x 8 with Ly = Lx ny / nx, Lz = Lx nz / nx
(fluidsim/solvers/ns3d/test_solver.py:30-43).  Such lengths run through the
engine's mixed-radix passes (radix 8, 4, 2, 3, 5 and a direct DFT for other
primes); the oracle is numpy/pocketfft, which takes any length.

Bars (written in each test): FP64 transforms within 1e-13 relative L2 of
numpy's rfftn / irfftn; FP64 tendencies within 1e-12 and FP64 states after
CFL steps within 1e-11 of the oracle with the same dt sequence; FP32 states
within 1e-5.
"""

import numpy as np
import pytest

from fluidsim_amd import _lib
from oracle.operators import Oper2D, Oper3D
from oracle.solvers import NS2D, NS3D, NS3DStrat

from .gpu_helpers import (host, make_plan, rel_l2, to_dev_complex, to_dev_real,
                          zeros_complex, zeros_real)

pytestmark = pytest.mark.gpu


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nh", [9, 12, 48, 6, 10, 15, 7, 2, 100])
def test_transforms_2d(nh, prec):
    """fsx_fft / fsx_ifft on nh x nh and nh x (nh + 3) grids against numpy."""
    for ny, nx in ((nh, nh), (nh + 3, nh)):
        print(f"case 2D transforms ny={ny} nx={nx} prec={prec}", flush=True)
        op = Oper2D(nx, ny, Lx=6.0, Ly=6.0 * ny / nx, coef_dealiasing=False)
        plan = make_plan(op, prec, 4, dealias=False)
        rng = np.random.default_rng(nx * 100 + ny)
        x = rng.standard_normal(op.shapeX)
        xd = to_dev_real(x, prec)
        kd = zeros_complex(op.shapeK, prec)
        _lib.call("fsx_fft", plan.handle, xd.data_ptr(), kd.data_ptr(), 1, _stream())
        ref = np.fft.rfft2(x, norm="forward")
        err = rel_l2(host(kd), ref)
        tol = 1e-13 if prec == 8 else 2e-6
        assert err < tol, (nx, ny, err)
        # inverse of a Hermitian half spectrum
        k = np.fft.rfft2(rng.standard_normal(op.shapeX), norm="forward")
        kd2 = to_dev_complex(k, prec)
        xo = zeros_real(op.shapeX, prec)
        _lib.call("fsx_ifft", plan.handle, kd2.data_ptr(), xo.data_ptr(), 1, _stream())
        ref = np.fft.irfft2(k, s=op.shapeX, norm="forward")
        err = rel_l2(host(xo), ref)
        assert err < tol, (nx, ny, err)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("shape", [(16, 12, 8), (9, 10, 6), (15, 12, 9), (48, 48, 48),
                                   (96, 96, 96), (20, 14, 11)])
def test_transforms_3d(shape, prec):
    nx, ny, nz = shape
    print(f"case 3D transforms {shape} prec={prec}", flush=True)
    op = Oper3D(nx, ny, nz, coef_dealiasing=False)
    plan = make_plan(op, prec, 7, dealias=False)
    rng = np.random.default_rng(nx + 7 * ny + 31 * nz)
    x = rng.standard_normal(op.shapeX)
    xd = to_dev_real(x, prec)
    kd = zeros_complex(op.shapeK, prec)
    _lib.call("fsx_fft", plan.handle, xd.data_ptr(), kd.data_ptr(), 1, _stream())
    tol = 1e-13 if prec == 8 else 2e-6
    err = rel_l2(host(kd), np.fft.rfftn(x, norm="forward"))
    assert err < tol, err
    k = np.fft.rfftn(rng.standard_normal(op.shapeX), norm="forward")
    kd2 = to_dev_complex(k, prec)
    xo = zeros_real(op.shapeX, prec)
    _lib.call("fsx_ifft", plan.handle, kd2.data_ptr(), xo.data_ptr(), 1, _stream())
    err = rel_l2(host(xo), np.fft.irfftn(k, s=op.shapeX, norm="forward"))
    assert err < tol, err


def test_default_params_ns2d_and_ns3d_run():
    """Simul(Simul.create_default_params()) on the reference default grid
    (48^2 and 48^3, constant init) builds and runs its default time loop."""
    import os
    import tempfile

    from fluidsim_amd.solvers.ns2d.solver import Simul as S2
    from fluidsim_amd.solvers.ns3d.solver import Simul as S3

    with tempfile.TemporaryDirectory() as d:
        old = os.environ.get("FLUIDSIM_PATH")
        os.environ["FLUIDSIM_PATH"] = d
        try:
            for Simul in (S2, S3):
                params = Simul.create_default_params()
                params.time_stepping.t_end = 1.0
                sim = Simul(params)
                assert sim.oper.nx == 48
                sim.time_stepping.start()
                assert sim.time_stepping.t >= 1.0
        finally:
            if old is None:
                del os.environ["FLUIDSIM_PATH"]
            else:
                os.environ["FLUIDSIM_PATH"] = old


def _ns2d_noise(op, kmax, seed):
    """band-limited solenoidal noise on any grid (oracle operators)"""
    rng = np.random.default_rng(seed)
    ux = rng.standard_normal(op.shapeK) + 1j * rng.standard_normal(op.shapeK)
    uy = rng.standard_normal(op.shapeK) + 1j * rng.standard_normal(op.shapeK)
    band = (op.K2 <= kmax**2) & (op.K2 > 0)
    ux[~band] = 0
    uy[~band] = 0
    op.projection_perp(ux, uy)
    ux = op.fft(op.ifft(ux))
    uy = op.fft(op.ifft(uy))
    rot = op.rotfft_from_vecfft(ux, uy)
    op.dealiasing(rot)
    return rot / np.abs(op.ifft(rot)).max()


@pytest.mark.parametrize("nh", [48, 9, 12, 30])
def test_ns2d_steps_non_pow2(nh):
    """ns2d FP64 RK4 with CFL 0.5, nu_2 = 1e-3, 8 steps: state within 1e-11
    of the oracle and the same dt sequence (nh x nh and nh x 2nh/3)."""
    from fluidsim_amd.solvers.ns2d.solver import Simul

    for nx, ny in ((nh, nh), (nh, max(2, 2 * nh // 3))):
        print(f"case ns2d steps nx={nx} ny={ny}", flush=True)
        Lx, Ly = 6.0, 6.0 * ny / nx
        op = Oper2D(nx, ny, Lx=Lx, Ly=Ly)
        rot0 = _ns2d_noise(op, kmax=3 * op.deltakx, seed=nx + ny)
        params = Simul.create_default_params()
        params.oper.nx, params.oper.ny = nx, ny
        params.oper.Lx, params.oper.Ly = Lx, Ly
        params.nu_2 = 1e-3
        params.time_stepping.USE_CFL = True
        params.time_stepping.cfl_coef = 0.5
        params.time_stepping.USE_T_END = False
        params.time_stepping.it_end = 8
        params.init_fields.type = "in_script"
        params.output.HAS_TO_SAVE = False
        params.output.periods_print.print_stdout = 0
        sim = Simul(params)
        sim.state.init_statespect_from(rot_fft=rot0)
        dts = []
        orig = sim.time_stepping.one_time_step_computation

        def record():
            orig()
            dts.append(sim.time_stepping.deltat)

        sim.time_stepping.one_time_step_computation = record
        sim.time_stepping.start()
        got = sim.state.get_var_numpy("rot_fft")
        orc = NS2D(op, nu_2=1e-3, USE_CFL=True, cfl_coef=0.5)
        orc.init_from_rotfft(rot0)
        odts = orc.run(8, record_dt=True)
        err = rel_l2(got, orc.state_spect[0])
        print(f"ns2d {nx}x{ny}: rel L2 = {err:.3e}")
        np.testing.assert_allclose(dts, odts, rtol=1e-12)
        assert err <= 1e-11


def _ns3d_noise(op, kmax, seed):
    rng = np.random.default_rng(seed)
    vv = []
    band = (op.K2 <= kmax**2) & (op.K2 > 0)
    for _ in range(3):
        a = rng.standard_normal(op.shapeK) + 1j * rng.standard_normal(op.shapeK)
        a[~band] = 0
        vv.append(a)
    op.project_perpk3d(*vv)
    vv = [op.fft(op.ifft(a)) for a in vv]
    for a in vv:
        op.dealiasing(a)
    vmax = max(np.abs(op.ifft(a)).max() for a in vv)
    return [a / vmax for a in vv]


def _ns3d_sim(nx, ny, nz, Lx, Ly, Lz, precision="float64", nsteps=5, strat=False, **extra):
    if strat:
        from fluidsim_amd.solvers.ns3d.strat.solver import Simul
    else:
        from fluidsim_amd.solvers.ns3d.solver import Simul
    params = Simul.create_default_params()
    po = params.oper
    po.nx, po.ny, po.nz = nx, ny, nz
    po.Lx, po.Ly, po.Lz = Lx, Ly, Lz
    po.precision = precision
    for k, v in extra.items():
        params[k] = v
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = nsteps
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    return Simul(params)


@pytest.mark.parametrize("shape", [(16, 12, 8), (15, 12, 9), (96, 96, 96), (24, 20, 18)])
def test_ns3d_steps_non_pow2(shape):
    """ns3d FP64 RK4, CFL 0.5, nu_2 = 1e-2, 5 steps on the reference's test
    grid shapes (box Lx = 6, Ly = Lx ny/nx, Lz = Lx nz/nx): state within
    1e-11 of the oracle, same dt sequence."""
    nx, ny, nz = shape
    print(f"case ns3d steps {shape}", flush=True)
    Lx = 6.0
    Ly, Lz = Lx * ny / nx, Lx * nz / nx
    op = Oper3D(nx, ny, nz, Lx=Lx, Ly=Ly, Lz=Lz)
    vv = _ns3d_noise(op, kmax=3 * op.deltakx, seed=nx)
    sim = _ns3d_sim(nx, ny, nz, Lx, Ly, Lz, nu_2=1e-2)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    dts = []
    orig = sim.time_stepping.one_time_step_computation

    def record():
        orig()
        dts.append(sim.time_stepping.deltat)

    sim.time_stepping.one_time_step_computation = record
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()
    orc = NS3D(op, nu_2=1e-2, USE_CFL=True, cfl_coef=0.5)
    orc.state_spect[:] = np.array(vv)
    odts = orc.run(5, record_dt=True)
    err = rel_l2(got, orc.state_spect)
    print(f"ns3d {shape}: rel L2 = {err:.3e}")
    np.testing.assert_allclose(dts, odts, rtol=1e-12)
    assert err <= 1e-11


def test_ns3d_fp32_non_pow2():
    """FP32 engine on 48^3 (5 CFL steps) within 1e-5 of the FP64 oracle."""
    n = 48
    op = Oper3D(n, n, n)
    vv = _ns3d_noise(op, kmax=4.0, seed=3)
    sim = _ns3d_sim(n, n, n, 2 * np.pi, 2 * np.pi, 2 * np.pi, precision="float32",
                    nu_2=1e-2)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy().astype(np.complex128)
    orc = NS3D(op, nu_2=1e-2, USE_CFL=True, cfl_coef=0.5)
    orc.state_spect[:] = np.array(vv)
    orc.run(5)
    err = rel_l2(got, orc.state_spect)
    print(f"ns3d FP32 48^3: rel L2 vs FP64 oracle = {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("shape", [(24, 18, 12), (15, 12, 9)])
def test_strat_steps_non_pow2(shape):
    """ns3d.strat FP64 (N = 2), 4 CFL steps with a buoyancy field: within
    1e-11 of the oracle."""
    nx, ny, nz = shape
    print(f"case strat steps {shape}", flush=True)
    Lx = 6.0
    Ly, Lz = Lx * ny / nx, Lx * nz / nx
    op = Oper3D(nx, ny, nz, Lx=Lx, Ly=Ly, Lz=Lz)
    vv = _ns3d_noise(op, kmax=3 * op.deltakx, seed=nz)
    b = 0.3 * _ns3d_noise(op, kmax=3 * op.deltakx, seed=nz + 1)[0]
    sim = _ns3d_sim(nx, ny, nz, Lx, Ly, Lz, strat=True, nsteps=4, nu_2=1e-2, N=2.0)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2], b_fft=b)
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()
    orc = NS3DStrat(op, N=2.0, nu_2=1e-2, USE_CFL=True, cfl_coef=0.5)
    orc.state_spect[:] = np.array(list(vv) + [b])
    orc.run(4)
    err = rel_l2(got, orc.state_spect)
    print(f"strat {shape}: rel L2 = {err:.3e}")
    assert err <= 1e-11
