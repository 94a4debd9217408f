"""Fused Runge-Kutta steps replayed from a HIP graph
(params.time_stepping.hip_graph) against the eager launches.

The graph holds the same launches in the same order, with every per-step
value in device memory (dt from the CFL rule, the forcing normalisation's
coefficients pushed before each replay), so the states must be equal bit for
bit after every step: each step ends with the host reading dt and the NaN
flag, and the test copies both states to the host between replays.  Cases:
a forced ns2d run (tcrandom, CFL, FP64), the learned-closure LES (FP32,
forced, network term in every stage), ns3d FP32 with CFL, and a forced ns2d
run whose outputs are due on some steps (those steps run eagerly, with the
CFL update and forcing ahead of the outputs; the saved spatial means agree
to 1e-12, the order of the diagnostics pass's FP64 LDS atomics varying from
run to run).
"""

import contextlib
import sys

import numpy as np
import pytest

from fluidsim_amd import closure as CL
from fluidsim_amd.synthetic import ns2d_noise_rot_fft, ns3d_spectrum_vv_fft

pytestmark = pytest.mark.gpu


def _ns2d_pair(n, precision, les=False, outputs=False):
    sims = []
    for graph in (False, True):
        Cls, p = CL.ns2d_params(n, 1e-3, precision=precision, nkmin=3, nkmax=5, les=les)
        p.time_stepping.hip_graph = graph
        if outputs:
            p.output.periods_save.spatial_means = 0.05
        np.random.seed(3)
        with contextlib.redirect_stdout(sys.stderr):
            sim = Cls(p)
        sim.state.init_statespect_from(rot_fft=ns2d_noise_rot_fft(n, n, seed=4, kmax=12))
        sims.append(sim)
    return sims


def _step_and_compare(sims, nsteps):
    import torch

    for it in range(nsteps):
        np.random.seed(100 + it)  # the forcing renewals draw from numpy's generator
        with contextlib.redirect_stdout(sys.stderr):
            sims[0].time_stepping.one_time_step()
        np.random.seed(100 + it)
        with contextlib.redirect_stdout(sys.stderr):
            sims[1].time_stepping.one_time_step()
        a = sims[0].state.state_spect.data.cpu()
        b = sims[1].state.state_spect.data.cpu()
        assert torch.equal(a, b), f"step {it}"
        ta, tb = sims[0].time_stepping, sims[1].time_stepping
        assert ta.deltat == tb.deltat and ta.t == tb.t, f"step {it}"


def _prepare(sims):
    with contextlib.redirect_stdout(sys.stderr):
        for s in sims:
            s.time_stepping.prepare_main_loop()


def test_graph_ns2d_forced_fp64():
    print("case graph step ns2d 64^2 FP64 forced CFL", flush=True)
    sims = _ns2d_pair(64, "float64")
    _prepare(sims)
    _step_and_compare(sims, 12)
    assert len(sims[1].time_stepping._graphs) == 1


def test_graph_les_closure_fp32():
    print("case graph step LES 64^2 FP32 with the network", flush=True)
    sims = _ns2d_pair(64, "float32", les=True)
    net = CL.CNNClosure(n_in=1, nlayers=6, seed=2)
    for s in sims:
        s.closure_net = net
    _prepare(sims)
    _step_and_compare(sims, 10)


def test_graph_ns2d_outputs_due():
    print("case graph step ns2d 64^2 FP64 forced, outputs due", flush=True)
    sims = _ns2d_pair(64, "float64", outputs=True)
    _prepare(sims)
    _step_and_compare(sims, 20)
    m0 = sims[0].output.spatial_means
    m1 = sims[1].output.spatial_means
    assert len(m0.records) == len(m1.records) > 1
    # the states are equal; the saved means come from the diagnostics pass,
    # whose FP64 LDS histogram atomics add in a run-dependent order
    for r0, r1 in zip(m0.records, m1.records):
        assert r0.keys() == r1.keys()
        for k in r0:
            assert np.allclose(np.asarray(r0[k]), np.asarray(r1[k]), rtol=1e-12, atol=1e-300), k


def test_graph_ns3d_fp32():
    from fluidsim_amd.solvers.ns3d.solver import Simul

    print("case graph step ns3d 32^3 FP32 CFL", flush=True)
    vv = ns3d_spectrum_vv_fft(32, seed=5, dtype=np.complex64)
    sims = []
    for graph in (False, True):
        p = Simul.create_default_params()
        p.oper.nx = p.oper.ny = p.oper.nz = 32
        p.oper.precision = "float32"
        p.nu_2 = 1e-3
        p.time_stepping.hip_graph = graph
        p.init_fields.type = "in_script"
        p.output.HAS_TO_SAVE = False
        p.output.periods_print.print_stdout = 0
        with contextlib.redirect_stdout(sys.stderr):
            sim = Simul(p)
        sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
        sims.append(sim)
    _prepare(sims)
    _step_and_compare(sims, 8)
