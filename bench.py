#!/usr/bin/env python
"""Benchmark: RK4 steps/s of the ns3d solver on one MI355X.

Workload (default): ns3d, periodic (2 pi)^3, 1024^3, FP32, RK4, order-8
hyperviscosity with k_max * eta = 1, seeded isotropic solenoidal initial
velocity with E(k) ~ k^4 exp(-2 (k/4)^2) normalised to unit energy (drawn on
the device at 1024^3), fixed dt taken from the initial CFL condition at 0.5.
A step is one full RK4 step (4 fused tendency evaluations + 4 Runge-Kutta
stages) with the state resident in HBM.  The same line carries the 512^3
FP32 and FP64 rates of the workload (``cfg2``), the FP32-vs-FP64 and
FP64-oracle accuracy (``accuracy``) and the CPU baseline.

``--ensemble``: the batched ns2d ensemble instead (512 independent FP32
runs: 256 at 128^2, 192 at 256^2, 64 at 512^2, seed = run index, nu_2
log-uniform in [1e-4, 1e-2], RK4 at each run's fixed initial-CFL dt, one
HIP graph launch per ensemble step); the value is ensemble steps/s.

``--closure``: the learned-closure workload instead: forced ns2d DNS
(tcrandom forcing at k in [8, 10], nu_2 = 1e-5) generates filtered
training pairs, the convolutional closure is trained on the device, then a
held-out DNS is compared with the coarse LES with and without the closure
over the same simulated time; the value is the DNS / LES wall-time ratio,
with the time-averaged energy-spectrum errors against the filtered DNS.

Prints ONE JSON line with the metric, a ``roofline`` object for the
dominant kernel (algorithmic bytes per launch / average launch duration,
measured with HIP events over the timed region), the step-level roofline of
the algorithmic bytes B_alg = T (b_X + b_K) + R b_K per RK4 step
(T = 36 real transforms, R = 51 state streams for ns3d), and a
``cpu_baseline`` from the FP64 numpy oracle timed on the host cores.
"""

import argparse
import contextlib
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "RK4 steps/s + %HBM roofline; closure-LES vs DNS speedup at equal error"
PEAK_HBM_GBPS = 8000.0
KINDS = ["z-pass inverse", "y-pass inverse", "x row pass", "y-pass forward",
         "z-pass forward", "RK stage", "other", "diagnostics pass"]


def b_alg_ns3d(n, prec):
    bX = prec * n**3
    bK = 2 * prec * n * n * (n // 2 + 1)
    return 36 * (bX + bK) + 51 * bK


def make_sim(n, precision, seed=0, vv=None, device_init=False):
    """ns3d n^3 sim of the workload; the initial field comes from vv, from the
    host generator, or (device_init, for 1024^3) from the device generator."""
    from fluidsim_amd.solvers.ns3d.solver import Simul
    from fluidsim_amd.synthetic import (ns3d_spectrum_vv_device, ns3d_spectrum_vv_fft,
                                        nu8_for_kmax_eta)

    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.oper.precision = precision
    params.nu_8 = nu8_for_kmax_eta(n)
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    sim = Simul(params)
    cdt = np.complex64 if precision == "float32" else np.complex128
    if vv is None and device_init:
        vv = ns3d_spectrum_vv_device(sim.oper, seed=seed)
        sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
        del vv
    else:
        if vv is None:
            vv = ns3d_spectrum_vv_fft(n, seed=seed)
        # the generator rounds its FP64 field at the end: a cast of the FP64
        # field is the FP32 field of the same seed
        sim.state.init_statespect_from(vx_fft=vv[0].astype(cdt), vy_fft=vv[1].astype(cdt),
                                       vz_fft=vv[2].astype(cdt))
    # fixed dt from the initial CFL condition
    sim.time_stepping.compute_time_increment_CLF()
    params.time_stepping.USE_CFL = False
    return sim


def b_alg_strat(n, prec):
    """T (b_X + b_K) + R b_K per RK4 step of ns3d.strat, T = 52, R = 68."""
    bX = prec * n**3
    bK = 2 * prec * n * n * (n // 2 + 1)
    return 52 * (bX + bK) + 68 * bK


# anisotropic forcing region of the strat workload (45 degrees, 10 degrees wide)
STRAT_ANISO = {"angle": np.pi / 4, "delta_angle": np.radians(10)}


def make_strat_sim(n, precision, seed=0, fixed_dt=None):
    """ns3d.strat, N = 4, Pr = 1 (one nu_8 for v and b), nu_8 with
    k_max eta = 1, toroidal time-correlated random forcing at k in [2, 4]
    (tcrandom_anisotropic, key_forced = vt_fft), isotropic velocity
    E(k) ~ k^4 exp(-2 (k/4)^2) (seed), b = 0; CFL 0.5, or a fixed dt."""
    from fluidsim_amd.solvers.ns3d.strat.solver import Simul
    from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft, nu8_for_kmax_eta

    params = Simul.create_default_params()
    po = params.oper
    po.nx = po.ny = po.nz = n
    po.Lx = po.Ly = po.Lz = 2 * np.pi
    po.precision = precision
    params.N = 4.0
    params.nu_8 = nu8_for_kmax_eta(n)
    pt = params.time_stepping
    pt.USE_CFL = fixed_dt is None
    pt.cfl_coef = 0.5
    if fixed_dt is not None:
        pt.deltat0 = fixed_dt
    pt.USE_T_END = False
    params.forcing.enable = True
    params.forcing.type = "tcrandom_anisotropic"
    params.forcing.key_forced = "vt_fft"
    params.forcing.nkmin_forcing = 2
    params.forcing.nkmax_forcing = 4
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    np.random.seed(seed)  # forcing seeds and renewals
    sim = Simul(params)
    cdt = np.complex64 if precision == "float32" else np.complex128
    if n > 512:
        # the device generator (the host one needs minutes at 1024^3)
        from fluidsim_amd.synthetic import ns3d_spectrum_vv_device

        vv = ns3d_spectrum_vv_device(sim.oper, seed=seed)
    else:
        vv = ns3d_spectrum_vv_fft(n, seed=seed, dtype=cdt)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    del vv
    return sim


def _energy_shells(oper, vv):
    from oracle.diagnostics import shell_spectrum

    return shell_spectrum(sum(0.5 * np.abs(a) ** 2 for a in vv), oper)


def _spectrum_error(E, E_ref, kmax):
    """relative L2 of an energy spectrum over shells 1 .. kmax (the
    dealiased range), and the largest relative error over the shells that
    hold at least 1e-6 of the peak (below that FP32 round-off of the
    largest modes dominates a shell)"""
    a, r = E[1: kmax + 1], E_ref[1: kmax + 1]
    keep = r >= 1e-6 * r.max()
    return {"rel_l2": float(np.linalg.norm(a - r) / np.linalg.norm(r)),
            "max_rel_shells_above_1e-6_of_peak": float(np.max(np.abs(a[keep] / r[keep] - 1))),
            "shells_above_1e-6_of_peak": int(keep.sum())}


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _device_rel_l2(torch, a, ref):
    """relative L2 of two device states (FP64 accumulation, field by field)"""
    num = den = 0.0
    for f in range(ref.shape[0]):
        r = ref[f]
        num += float(torch.linalg.vector_norm(a[f].to(r.dtype) - r).item() ** 2)
        den += float(torch.linalg.vector_norm(r).item() ** 2)
    return (num / den) ** 0.5


def ns3d_accuracy(torch, n_oracle=128, steps_oracle=5, seed=0):
    """The FP64 and FP32 engines against the FP64 oracle at n_oracle: states
    after steps_oracle fixed-dt steps on the same seeded fields (relative L2)
    and their shell energy spectra."""
    from oracle import operators as O
    from oracle.solvers import NS3D
    from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft, nu8_for_kmax_eta

    res = {}
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    O.set_workers(cores)
    op = O.Oper3D(n_oracle, n_oracle, n_oracle)
    v0 = ns3d_spectrum_vv_fft(n_oracle, seed=seed)
    s64 = make_sim(n_oracle, "float64", seed, vv=v0)
    dt = s64.time_stepping.deltat
    orc = NS3D(op, nu_8=nu8_for_kmax_eta(n_oracle), deltat0=dt)
    orc.state_spect[:] = np.array(v0)
    orc.run(steps_oracle)
    ref = orc.state_spect
    E_ref = _energy_shells(op, ref)
    for prec in ("float64", "float32"):
        sim = s64 if prec == "float64" else make_sim(n_oracle, prec, seed, vv=v0)
        sim.time_stepping.deltat = dt
        for _ in range(steps_oracle):
            sim.time_stepping.one_time_step_computation()
        got = sim.state.state_spect.to_numpy().astype(np.complex128)
        res[f"{prec}_vs_oracle"] = {
            "state_rel_l2": _rel_l2(got, ref),
            "energy_spectrum": _spectrum_error(_energy_shells(op, got), E_ref,
                                               n_oracle // 3)}
        del sim
    res["oracle_case"] = (f"ns3d {n_oracle}^3, {steps_oracle} RK4 steps at the FP64 "
                          f"initial-CFL dt, same seeded fields, FP64 numpy oracle")
    del s64
    gc.collect()
    return res


def _timed_steps(torch, ts, nsteps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ts.one_time_step_computation()
        ts.t += ts.deltat
        ts.it += 1
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def cfg2_legs(torch, n=512, warmup=2, steps=10, steps_total=100, seed=0):
    """The 512^3 workload in FP32 and FP64 on the same seeded fields and the
    same dt: RK4 steps/s of each (warmup, then `steps` timed steps with the
    state in HBM) and its step roofline; then both continue to steps_total
    steps and the FP32 state and shell energy spectrum are compared with the
    FP64 engine's on the device (the FP64 engine holds the oracle to 1e-11
    at every size the tests run)."""
    from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft

    vv = ns3d_spectrum_vv_fft(n, seed=seed)
    with contextlib.redirect_stdout(sys.stderr):
        s64 = make_sim(n, "float64", seed, vv=vv)
        s32 = make_sim(n, "float32", seed, vv=vv)
    del vv
    s32.time_stepping.deltat = s64.time_stepping.deltat
    res = {"case": (f"ns3d {n}^3 RK4, same workload and seeded fields as the headline "
                    f"(host generator), same fixed dt in both precisions")}
    for prec, sim in (("float32", s32), ("float64", s64)):
        ts = sim.time_stepping
        for _ in range(warmup):
            ts.one_time_step_computation()
        el = _timed_steps(torch, ts, steps)
        sps = steps / el
        balg = b_alg_ns3d(n, 4 if prec == "float32" else 8)
        res[prec] = {"steps_per_s": sps, "ms_per_step": el / steps * 1e3,
                     "steps": steps, "warmup": warmup,
                     "step_roofline": {"B_alg_GB_per_step": balg / 1e9,
                                       "achieved_GBps": balg * sps / 1e9,
                                       "frac": balg * sps / 1e9 / PEAK_HBM_GBPS}}
    for sim in (s32, s64):
        for _ in range(steps_total - warmup - steps):
            sim.time_stepping.one_time_step_computation()
    torch.cuda.synchronize()
    E = [sim.output.spectra.compute(kzkh=False)[1]["spectra_E"] for sim in (s32, s64)]
    res["float32_vs_float64_engine"] = {
        "case": f"ns3d {n}^3, {steps_total} RK4 steps at the same dt",
        "state_rel_l2": _device_rel_l2(torch, s32.state.state_spect.data,
                                       s64.state.state_spect.data),
        "energy_spectrum": _spectrum_error(E[0], E[1], n // 3)}
    del s32, s64
    gc.collect()
    torch.cuda.empty_cache()
    return res


def strat_oracle_check(n=256, steps=3, seed=0):
    """FP64 and FP32 engine states of the strat workload at n^3 after
    `steps` forced steps at the FP64 initial-CFL dt against the FP64 oracle
    (same forcing draws from numpy's global generator); the oracle's own
    time per step is the CPU baseline."""
    from oracle import operators as O
    from oracle.forcing import TCRandomForcing
    from oracle.solvers import NS3DStrat
    from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft, nu8_for_kmax_eta

    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    O.set_workers(cores)
    s64 = make_strat_sim(n, "float64", seed)
    s64.time_stepping.compute_time_increment_CLF()
    dt = s64.time_stepping.deltat
    del s64
    got = {}
    for prec in ("float64", "float32"):
        sim = make_strat_sim(n, prec, seed, fixed_dt=dt)
        sim.time_stepping.prepare_main_loop()
        for _ in range(steps):
            sim.time_stepping.one_time_step()
        got[prec] = sim.state.state_spect.to_numpy().astype(np.complex128)
        del sim
    np.random.seed(seed)
    op = O.Oper3D(n, n, n)
    orc = NS3DStrat(op, N=4.0, nu_8=nu8_for_kmax_eta(n), deltat0=dt)
    vv = ns3d_spectrum_vv_fft(n, seed=seed)
    orc.state_spect[:] = np.array(list(vv) + [np.zeros_like(vv[0])])
    orc.forcing = TCRandomForcing(orc.oper, NS3DStrat.keys_state_spect, 2, 4,
                                  key_forced="vt_fft", anisotropic=STRAT_ANISO)
    t0 = time.perf_counter()
    orc.run(steps)
    t_step = (time.perf_counter() - t0) / steps
    ref = orc.state_spect
    E_ref = _energy_shells(op, ref[:3])
    res = {"case": (f"ns3d.strat {n}^3 forced, {steps} RK4 steps at dt = {dt:.4g}, "
                    f"same seeded fields and forcing draws, FP64 numpy oracle")}
    for prec, a in got.items():
        res[f"{prec}_vs_oracle"] = {
            "state_rel_l2": _rel_l2(a, ref),
            "energy_spectrum": _spectrum_error(_energy_shells(op, a[:3]), E_ref, n // 3)}
    return res, t_step, cores


def bench_ns2d(args, torch, _lib, copy_gbps):
    """ns2d at n^2 (the DNS of the closure workload at 2048^2 FP32): forced
    (tcrandom at k in [8, 10]), nu_2 = 1e-5, CFL 0.5; steps/s with the state
    in HBM, step roofline (B_alg = 21 (b_X + b_K) + 17 b_K), per-pass table."""
    from fluidsim_amd import closure as CL
    from fluidsim_amd.synthetic import ns2d_noise_rot_fft

    n, prec = args.n, (4 if args.precision == "float32" else 8)
    Cls, params = CL.ns2d_params(n, 1e-5, precision=args.precision)
    np.random.seed(0)
    with contextlib.redirect_stdout(sys.stderr):
        sim = Cls(params)
        sim.state.init_statespect_from(
            rot_fft=ns2d_noise_rot_fft(n, n, seed=0, kmax=12, velo_max=0.5))
        ts = sim.time_stepping
        ts.prepare_main_loop()
    plan = sim.oper.plan.handle
    for _ in range(args.warmup):
        ts.one_time_step()
    torch.cuda.synchronize()
    _lib.call("fsx_timing_enable", plan, 1)
    ms8 = (_lib.ctypes.c_double * 8)()
    by8 = (_lib.ctypes.c_double * 8)()
    cn8 = (_lib.ctypes.c_longlong * 8)()
    _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
    _lib.call("fsx_timing_enable", plan, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts.one_time_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # per-pass durations from a second, event-timed run of the same steps
    _lib.call("fsx_timing_enable", plan, 1)
    for _ in range(args.steps):
        ts.one_time_step()
    _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
    _lib.call("fsx_timing_enable", plan, 0)
    steps_per_s = args.steps / elapsed
    # the same run with its steps replayed from a HIP graph
    del sim, ts
    gc.collect()
    Cls, params = CL.ns2d_params(n, 1e-5, precision=args.precision, graph=True)
    np.random.seed(0)
    with contextlib.redirect_stdout(sys.stderr):
        sim = Cls(params)
        sim.state.init_statespect_from(
            rot_fft=ns2d_noise_rot_fft(n, n, seed=0, kmax=12, velo_max=0.5))
        ts = sim.time_stepping
        ts.prepare_main_loop()
    for _ in range(max(args.warmup, 3)):
        ts.one_time_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts.one_time_step()
    torch.cuda.synchronize()
    graph_steps_per_s = args.steps / (time.perf_counter() - t0)
    # in 2D the last forward pass (timing slot 4) is the y-pass with the
    # Runge-Kutta stage
    names = list(KINDS)
    names[4] = "y-pass forward + RK stage"
    kernels = {}
    for k in range(7):
        if cn8[k]:
            kernels[names[k]] = {
                "avg_ms": ms8[k] / cn8[k],
                "launches_per_step": cn8[k] / args.steps,
                "alg_GBps": by8[k] / (ms8[k] * 1e-3) / 1e9,
            }
    dom = max(range(7), key=lambda k: ms8[k])
    achieved = by8[dom] / (ms8[dom] * 1e-3) / 1e9
    balg = b_alg_ns2d(n, prec)
    result = {
        "metric": METRIC, "value": steps_per_s, "unit": "RK4 steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if prec == 4 else "f64",
        "data": "synthetic",
        "config": {"workload": (f"ns2d {n}^2 {args.precision} RK4, nu_2 = 1e-5, tcrandom "
                                "forcing at k in [8, 10], CFL 0.5 (the closure workload's "
                                "DNS)"),
                   "grid": [n, n], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": achieved,
                     "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBPS,
                     "traffic": None, "copy_peak_measured": copy_gbps},
        "step_roofline": {"B_alg_GB_per_step": balg / 1e9,
                          "achieved_GBps": balg * steps_per_s / 1e9,
                          "frac": balg * steps_per_s / 1e9 / PEAK_HBM_GBPS},
        "hip_graph": {"steps_per_s": graph_steps_per_s,
                      "step_roofline_frac": balg * graph_steps_per_s / 1e9 / PEAK_HBM_GBPS},
        "kernels": kernels,
    }
    return result


def bench_strat(args, torch, _lib, copy_gbps):
    """ns3d.strat workload at n^3 FP32: steps/s with the state in HBM, step
    roofline, per-pass table, FP64-oracle error at 256^3, CPU baseline."""
    n, prec = args.n, (4 if args.precision == "float32" else 8)
    print(f"[bench] building ns3d.strat {n}^3 {args.precision}", file=sys.stderr, flush=True)
    with contextlib.redirect_stdout(sys.stderr):
        sim = make_strat_sim(n, args.precision, seed=0)
        ts = sim.time_stepping
        ts.prepare_main_loop()
    plan = sim.oper.plan.handle
    for _ in range(args.warmup):
        ts.one_time_step()
    torch.cuda.synchronize()
    print(f"[bench] warm-up done, dt = {ts.deltat:.4g}", file=sys.stderr, flush=True)
    _lib.call("fsx_timing_enable", plan, 1)
    ms8 = (_lib.ctypes.c_double * 8)()
    by8 = (_lib.ctypes.c_double * 8)()
    cn8 = (_lib.ctypes.c_longlong * 8)()
    _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts.one_time_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
    _lib.call("fsx_timing_enable", plan, 0)
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    steps_per_s = args.steps / elapsed
    kernels = {}
    for k in range(7):
        if cn8[k]:
            kernels[KINDS[k]] = {
                "avg_ms": ms8[k] / cn8[k],
                "launches_per_step": cn8[k] / args.steps,
                "share_of_step": ms8[k] / (elapsed * 1e3),
                "alg_GBps": by8[k] / (ms8[k] * 1e-3) / 1e9,
            }
    dom = max(range(7), key=lambda k: ms8[k])
    achieved = by8[dom] / (ms8[dom] * 1e-3) / 1e9
    balg = b_alg_strat(n, prec)
    t_final = ts.t
    del sim, ts
    gc.collect()
    torch.cuda.empty_cache()
    result = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == 4 else "f64",
        "data": "synthetic",
        "config": {
            "workload": (
                f"ns3d.strat {n}^3 {args.precision} RK4, N = 4, Pr = 1, nu_8 with "
                "k_max*eta = 1, toroidal tcrandom_anisotropic forcing at k in [2, 4], "
                "isotropic E(k)~k^4 exp(-2(k/4)^2) velocity (seed 0, "
                + ("device generator" if n > 512 else "host generator")
                + "), b = 0, CFL 0.5"
            ),
            "grid": [n, n, n],
            "parallelism": "single GPU",
            "t_after_timed_steps": t_final,
        },
        "roofline": {
            "bound": "hbm", "kernel": KINDS[dom], "achieved": achieved,
            "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBPS,
            "traffic": read_traffic(f"strat_{n}_{args.precision}:{KINDS[dom]}"),
            "copy_peak_measured": copy_gbps, "frac_of_copy": achieved / copy_gbps,
        },
        "step_roofline": {
            "B_alg_GB_per_step": balg / 1e9,
            "achieved_GBps": balg * steps_per_s / 1e9,
            "frac": balg * steps_per_s / 1e9 / PEAK_HBM_GBPS,
            "frac_of_copy": balg * steps_per_s / 1e9 / copy_gbps,
        },
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        print("[bench] FP64 oracle check and CPU baseline at "
              f"{args.cpu_n}^3", file=sys.stderr, flush=True)
        with contextlib.redirect_stdout(sys.stderr):
            acc, t_step, cores = strat_oracle_check(args.cpu_n, steps=args.oracle_steps)
        result["accuracy"] = acc
        m = args.cpu_n
        scale = (n**3 * math.log2(n**3)) / (m**3 * math.log2(m**3))
        result["cpu_baseline"] = {
            "value": 1.0 / (t_step * scale), "unit": "RK4 steps/s", "cores": cores,
            "kind": "port",
            "sample": (f"{args.oracle_steps} forced RK4 steps of the same workload at "
                       f"{m}^3 FP64 with the oracle ({t_step:.2f} s/step, scipy.fft "
                       f"workers={cores}), scaled by N log N (x{scale:.1f}) to {n}^3"),
        }
    return result


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_step_time(n, workers, steps=3):
    """Median wall time of `steps` RK4 steps of the ns3d workload at n^3 with
    the FP64 oracle (scipy.fft workers, numpy elementwise), after one
    warm-up step (fluidsim-bench style: fixed dt, no output)."""
    from oracle import operators as O
    from oracle.solvers import NS3D
    from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft, nu8_for_kmax_eta

    O.set_workers(workers)
    op = O.Oper3D(n, n, n)
    sim = NS3D(op, nu_8=nu8_for_kmax_eta(n), USE_CFL=True, cfl_coef=0.5)
    vv = ns3d_spectrum_vv_fft(n, seed=0)
    sim.state_spect[:] = np.array(vv)
    del vv
    sim.compute_time_increment_CLF()
    sim.USE_CFL = False
    sim.one_time_step_computation()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        sim.one_time_step_computation()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(n_metric, n_all=256, n_one=128):
    """The FP64 oracle on the host cores: all cores of this process (scipy.fft
    workers = cores, the analogue of mpirun -np cores) at n_all^3 and one
    core (fluidsim's default sequential mode) at n_one^3, each scaled by
    N log N to n_metric^3."""
    import scipy

    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    N2 = n_metric**3

    def scaled(n, workers):
        med, ts = _oracle_step_time(n, workers)
        N1 = n**3
        scale = (N2 * math.log2(N2)) / (N1 * math.log2(N1))
        return med, ts, scale

    med, ts, scale = scaled(n_all, cores)
    med1, ts1, scale1 = scaled(n_one, 1)
    return {
        "value": 1.0 / (med * scale),
        "unit": "RK4 steps/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"median of {len(ts)} RK4 steps (after 1 warm-up) of the same ns3d "
                   f"workload at {n_all}^3 FP64 with the oracle: {med:.2f} s/step "
                   f"(scipy.fft workers={cores}, numpy elementwise single-threaded), "
                   f"scaled by N log N (x{scale:.1f}) to {n_metric}^3"),
        "cpu_model": _cpu_model(),
        "numpy": np.__version__,
        "scipy": scipy.__version__,
        "single_core": {
            "value": 1.0 / (med1 * scale1),
            "cores": 1,
            "sample": (f"median of {len(ts1)} RK4 steps (after 1 warm-up) at {n_one}^3 "
                       f"FP64, one core: {med1:.2f} s/step, scaled by N log N "
                       f"(x{scale1:.0f}) to {n_metric}^3"),
        },
    }


def measured_copy_bandwidth(torch, nbytes=1 << 30, reps=5):
    """Device-to-device copy of 1 GiB (read + write), timed with HIP events:
    the practical HBM ceiling the kernels are also compared with."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def time_diagnostics(sim, plan, torch, reps=3):
    """One save of each on-device output at the bench state (spectra with
    kz-kh, spectral energy budget, spatial means): wall time per call and
    the fsx_diag binning pass (live HIP events, engine timing slot 7),
    whose algorithmic bytes are the input fields read once plus the block
    records written."""
    from fluidsim_amd import _lib

    out = sim.output
    jobs = (("spectra", lambda: out.spectra.compute(kzkh=True)),
            ("spect_energy_budg", out.spect_energy_budg.compute),
            ("spatial_means", out.spatial_means.compute))
    ms8 = (_lib.ctypes.c_double * 8)()
    by8 = (_lib.ctypes.c_double * 8)()
    cn8 = (_lib.ctypes.c_longlong * 8)()
    res = {}
    for name, fn in jobs:
        fn()
        _lib.call("fsx_timing_enable", plan, 1)
        _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
        _lib.call("fsx_timing_enable", plan, 0)
        gbps = by8[7] / (ms8[7] * 1e-3) / 1e9 if ms8[7] else None
        res[name] = {
            "ms_per_call": wall * 1e3,
            "diag_pass_ms_per_call": ms8[7] / reps,
            "diag_pass_alg_GBps": gbps,
            "diag_pass_frac": gbps / PEAK_HBM_GBPS if gbps else None,
        }
    return res


ENSEMBLE_SPEC = ((128, 256), (256, 192), (512, 64))


def b_alg_ns2d(n, prec):
    """T (b_X + b_K) + R b_K per RK4 step of one ns2d run, T = 21, R = 17."""
    bX = prec * n * n
    bK = 2 * prec * n * (n // 2 + 1)
    return 21 * (bX + bK) + 17 * bK


def ensemble_runs(spec=ENSEMBLE_SPEC, seed0=0):
    from fluidsim_amd.synthetic import ns2d_noise_rot_fft

    rng = np.random.default_rng(seed0)
    runs, k = [], 0
    for n, count in spec:
        for _ in range(count):
            runs.append({"n": n, "nu_2": float(10 ** rng.uniform(-4, -2)),
                         "rot_fft": ns2d_noise_rot_fft(n, n, seed=seed0 + k, kmax=n // 8)})
            k += 1
    return runs


def cpu_baseline_ensemble(runs, steps=2):
    """FP64 oracle: a few RK4 steps of the first run of each size, scaled
    by the number of runs of that size."""
    from oracle import operators as O
    from oracle.solvers import NS2D

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    O.set_workers(cores)
    total, parts = 0.0, []
    for n, count in ENSEMBLE_SPEC:
        r = next(r for r in runs if r["n"] == n)
        sim = NS2D(O.Oper2D(n, n), nu_2=r["nu_2"], deltat0=r["deltat"])
        sim.init_from_rotfft(r["rot_fft"])
        sim.run(1)
        t0 = time.perf_counter()
        sim.run(steps)
        sec = (time.perf_counter() - t0) / steps
        total += sec * count
        parts.append(f"{n}^2: {sec * 1e3:.2f} ms/step x {count}")
    return {
        "value": 1.0 / total, "unit": "ensemble steps/s", "cores": cores, "kind": "port",
        "sample": f"{steps} RK4 steps of one run per size with the FP64 oracle "
                  f"(scipy.fft workers={cores}), scaled by the run counts: " + ", ".join(parts),
    }


def bench_ensemble(args, torch, _lib, copy_gbps):
    from fluidsim_amd.ensemble import EnsembleNS2D

    print("[bench] building the ensemble (512 runs)", file=sys.stderr, flush=True)
    runs = ensemble_runs()
    with contextlib.redirect_stdout(sys.stderr):
        ens = EnsembleNS2D(runs, precision="float32")
    ens.step(max(args.warmup, 1))
    ens.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(args.steps)
    ens.synchronize()
    elapsed = time.perf_counter() - t0
    ens.check_nan()
    steps_per_s = args.steps / elapsed
    balg = sum(b_alg_ns2d(r["n"], 4) for r in runs)
    # per-pass table from eager steps with live HIP events on every group
    ms8 = [0.0] * 8
    by8 = [0.0] * 8
    cn8 = [0] * 8
    m = (_lib.ctypes.c_double * 8)()
    b = (_lib.ctypes.c_double * 8)()
    c = (_lib.ctypes.c_longlong * 8)()
    for g in ens.groups:
        _lib.call("fsx_timing_enable", g.handle, 1)
        _lib.call("fsx_timing_read", g.handle, m, b, c, 1)
    ens.step(3, graph=False, serial=True)
    ens.synchronize()
    for g in ens.groups:
        _lib.call("fsx_timing_read", g.handle, m, b, c, 1)
        _lib.call("fsx_timing_enable", g.handle, 0)
        for k in range(8):
            ms8[k] += m[k]
            by8[k] += b[k]
            cn8[k] += c[k]
    # the 2D passes: y inverse, row pass, y forward with the RK stage
    names = dict(enumerate(KINDS))
    names[4] = "y-pass forward + RK stage"
    kernels = {}
    for k in range(7):
        if cn8[k]:
            kernels[names[k]] = {"ms_per_step": ms8[k] / 3,
                                 "alg_GBps": by8[k] / (ms8[k] * 1e-3) / 1e9,
                                 "launches_per_step": cn8[k] / 3}
    dom = max(range(7), key=lambda k: ms8[k])
    achieved = by8[dom] / (ms8[dom] * 1e-3) / 1e9
    result = {
        "metric": "ensemble steps/s (512 ns2d runs) + %HBM roofline",
        "value": steps_per_s,
        "unit": "ensemble RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "ns2d ensemble: 512 independent FP32 runs (256 at 128^2, 192 at "
                        "256^2, 64 at 512^2), seed = run index, nu_2 log-uniform in "
                        "[1e-4, 1e-2], RK4 at each run's initial-CFL(0.5) dt, 2/3 "
                        "dealiasing, one HIP graph launch per ensemble step",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": names[dom] + " (batched)",
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": None,
            "copy_peak_measured": copy_gbps,
            "frac_of_copy": achieved / copy_gbps,
        },
        "step_roofline": {
            "B_alg_GB_per_step": balg / 1e9,
            "achieved_GBps": balg * steps_per_s / 1e9,
            "frac": balg * steps_per_s / 1e9 / PEAK_HBM_GBPS,
            "frac_of_copy": balg * steps_per_s / 1e9 / copy_gbps,
        },
        "kernels": kernels,
        "tuning": {str(k): {"lines": v["lines"], "row_pass": v["row_pass"]}
                   for k, v in ens.tuning.items()},
    }
    runs = ens.runs  # with each run's dt
    ens.close()
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        with contextlib.redirect_stdout(sys.stderr):
            result["cpu_baseline"] = cpu_baseline_ensemble(runs)
    return result


def bench_closure(args, torch, _lib):
    from fluidsim_amd import closure as CL
    from fluidsim_amd.synthetic import ns2d_noise_rot_fft

    nd, nl, nu = args.closure_dns, args.closure_les, 1e-5

    def log(msg):
        print(f"[closure] {msg}", file=sys.stderr, flush=True)

    def new_dns(seed):
        Cls, p = CL.ns2d_params(nd, nu)
        np.random.seed(seed)
        with contextlib.redirect_stdout(sys.stderr):
            sim = Cls(p)
        sim.state.init_statespect_from(
            rot_fft=ns2d_noise_rot_fft(nd, nd, seed=seed, kmax=12, velo_max=0.5))
        return sim

    def run_until(sim, t):
        # the solver's messages go to stderr: stdout is the JSON line
        with contextlib.redirect_stdout(sys.stderr):
            return CL.run_until(sim, t)

    Cls, pl = CL.ns2d_params(nl, nu, forcing=False)
    with contextlib.redirect_stdout(sys.stderr):
        les_oper = Cls(pl).oper
    # 1. training pairs from forced DNS runs
    t0 = time.perf_counter()
    ws, ss = [], []
    for r in range(args.closure_runs):
        dns = new_dns(100 + r)
        run_until(dns, args.closure_spinup)
        for k in range(args.closure_snaps):
            run_until(dns, args.closure_spinup + (k + 1) * args.closure_snap_dt)
            w, sgs = CL.subgrid_terms(dns, les_oper)
            ws.append(w.float())
            ss.append(sgs.float())
        log(f"run {r}: t = {dns.time_stepping.t:.2f}, it = {dns.time_stepping.it}")
        del dns
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    # 2. training on the device
    net = CL.CNNClosure(n_in=1, nlayers=6, seed=0)
    t0 = time.perf_counter()
    half = args.closure_train_steps // 2
    sw, s_s, l1 = CL.train(net, ws, ss, half, batch=32, lr=1e-3, seed=0, log=100)
    _, _, l2 = CL.train(net, ws, ss, args.closure_train_steps - half, batch=32, lr=3e-4,
                        seed=1, log=100)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    log(f"training losses {l1[0]} ... {l2[-1]}")
    del ws, ss
    # 3. held-out DNS against the coarse LES with and without the closure
    dns = new_dns(999)
    run_until(dns, args.closure_spinup)
    wk = CL.cut_to_les(dns.oper, les_oper, dns.state.state_spect.data[0]).cpu().numpy()
    runs = {}
    for name in ("les_closure", "les_plain"):
        SimulLES, lp = CL.ns2d_params(nl, nu, les=True)
        np.random.seed(7)
        with contextlib.redirect_stdout(sys.stderr):
            les = SimulLES(lp)
        les.state.init_statespect_from(rot_fft=wk)
        les.time_stepping.t = dns.time_stepping.t
        if name == "les_closure":
            les.closure_net = net.folded(sw, s_s)
        runs[name] = les
    runs["dns"] = dns
    # same forcing in the three runs; each draws its renewals from its own
    # copy of the random state
    CL.share_forcing(dns, [runs["les_closure"], runs["les_plain"]])
    with contextlib.redirect_stdout(sys.stderr):
        spectra, wall, steps = CL.compare_runs(runs, les_oper, args.closure_eval_t,
                                               args.closure_sample_dt)
    rng_state = {"les_closure": np.random.get_state()}
    # per-launch rate of the hidden-layer convolutions (HIP events on the
    # LES stream) over extra LES steps after the timed evaluation
    les = runs["les_closure"]
    plan = les.oper.plan.handle
    c = _lib.ctypes
    ms, fl, cnt = c.c_double(), c.c_double(), c.c_longlong()
    _lib.call("fsx_timing_enable", plan, 1)
    _lib.call("fsx_timing_read_network", plan, ms, fl, cnt, 1)
    np.random.set_state(rng_state["les_closure"])
    for _ in range(20):
        les.time_stepping.one_time_step()
    _lib.call("fsx_timing_read_network", plan, ms, fl, cnt, 1)
    _lib.call("fsx_timing_enable", plan, 0)
    tflops = fl.value / (ms.value * 1e-3) / 1e12
    roofline = {
        "bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
        "frac": tflops / 2500.0, "traffic": None,
        "kernel": f"k_conv_fwd<32, 1>: hidden 32 -> 32 layer on one {nl}^2 field",
        "launch_us": ms.value / max(cnt.value, 1) * 1e3,
        "flops_per_launch": fl.value / max(cnt.value, 1),
        "launches": cnt.value,
    }
    kmax = int(nl / 3)  # kept shells of the LES grid (2/3 truncation)
    mean = {k: np.mean(v, axis=0)[1: kmax + 1] for k, v in spectra.items()}

    def err(k, lo=0):
        d = mean[k][lo:] - mean["dns"][lo:]
        return float(np.linalg.norm(d) / np.linalg.norm(mean["dns"][lo:]))

    def log_err(k):
        return float(np.mean(np.abs(np.log(mean[k] / mean["dns"]))))

    speedup = wall["dns"] / wall["les_closure"]
    result = {
        "metric": "closure-LES vs DNS speedup at equal error",
        "value": speedup,
        "unit": "DNS wall time / LES wall time",
        "n_gpus": 1,
        "steps": steps["les_closure"],
        "warmup": 0,
        "ms_per_step": wall["les_closure"] / max(steps["les_closure"], 1) * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 (network bf16 x bf16 -> f32)",
        "data": "synthetic",
        "config": {
            "workload": (f"ns2d forced (tcrandom k in [8, 10], rate 1), nu_2 = 1e-5; DNS "
                         f"{nd}^2 FP32; LES {nl}^2 FP32 with a 6-layer 5x5 32-channel GELU "
                         f"network (BF16 weights, FP32 accumulation) trained on "
                         f"{args.closure_runs} DNS runs x {args.closure_snaps} filtered "
                         f"snapshots; evaluation over {args.closure_eval_t} time units "
                         f"from a held-out DNS state"),
            "parallelism": "single GPU",
        },
        "spectrum_error": {
            "les_closure": err("les_closure"),
            "les_plain": err("les_plain"),
            "shells": [1, kmax],
            "definition": "relative L2 of the time-averaged energy spectrum vs the "
                          "filtered DNS",
            "small_scales": {
                "shells": [11, kmax],
                "les_closure": err("les_closure", 10),
                "les_plain": err("les_plain", 10),
            },
            "mean_abs_log_ratio": {
                "les_closure": log_err("les_closure"),
                "les_plain": log_err("les_plain"),
            },
        },
        "roofline": roofline,
        "wall_s": wall,
        "steps_by_run": steps,
        "ms_per_step_by_run": {k: wall[k] / max(steps[k], 1) * 1e3 for k in wall},
        "timings_s": {"training_data": t_data, "training": t_train},
        "training_loss": {"first": l1[0][1], "last": l2[-1][1]},
    }
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle)")
        result["cpu_baseline"] = cpu_baseline_closure(net.folded(sw, s_s), nd, nl, nu)
    return result


def closure_reduced(torch, nd=512, nl=64, nu=2e-4, band=(4, 6), nruns=4, nsnaps=50,
                    train_steps=1000, t_eval=4.0):
    """The learned-closure workload at reduced size (the GPU test's case):
    forced DNS nd^2 -> training pairs on nl^2 -> 6-layer network trained on
    the device -> held-out DNS against the LES with and without the network,
    same forcing.  Returns the
    DNS / LES wall-time ratio over the evaluation, the spectrum errors and
    the a priori error of the subgrid term."""
    from fluidsim_amd import closure as CL
    from fluidsim_amd.synthetic import ns2d_noise_rot_fft

    t00 = time.perf_counter()

    def new_dns(seed):
        Cls, p = CL.ns2d_params(nd, nu, nkmin=band[0], nkmax=band[1])
        np.random.seed(seed)
        sim = Cls(p)
        sim.state.init_statespect_from(
            rot_fft=ns2d_noise_rot_fft(nd, nd, seed=seed, kmax=12, velo_max=0.5))
        return sim

    with contextlib.redirect_stdout(sys.stderr):
        Cls, pl = CL.ns2d_params(nl, nu, forcing=False)
        les_oper = Cls(pl).oper
        ws, ss = [], []
        for r in range(nruns):
            dns = new_dns(100 + r)
            CL.run_until(dns, 3.0)
            for k in range(nsnaps):
                CL.run_until(dns, 3.0 + (k + 1) * 0.05)
                w, sg = CL.subgrid_terms(dns, les_oper)
                ws.append(w.float())
                ss.append(sg.float())
            del dns
        t_data = time.perf_counter() - t00
        net = CL.CNNClosure(n_in=1, nlayers=6, seed=0)
        t0 = time.perf_counter()
        sw, s_s, l1 = CL.train(net, ws, ss, train_steps // 2, batch=32, lr=1e-3, seed=0, log=100)
        _, _, l2 = CL.train(net, ws, ss, train_steps - train_steps // 2, batch=32, lr=3e-4,
                            seed=1, log=100)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        folded = net.folded(sw, s_s)
        dns = new_dns(999)
        CL.run_until(dns, 3.0)
        wk = CL.cut_to_les(dns.oper, les_oper, dns.state.state_spect.data[0]).cpu().numpy()
        runs = {}
        for name in ("les_closure", "les_plain"):
            SimulLES, lp = CL.ns2d_params(nl, nu, nkmin=band[0], nkmax=band[1], les=True)
            np.random.seed(7)
            les = SimulLES(lp)
            les.state.init_statespect_from(rot_fft=wk)
            les.time_stepping.t = dns.time_stepping.t
            if name == "les_closure":
                les.closure_net = folded
            runs[name] = les
        runs["dns"] = dns
        CL.share_forcing(dns, [runs["les_closure"], runs["les_plain"]])
        spectra, wall, steps = CL.compare_runs(runs, les_oper, t_eval, 0.2)
        kmax = nl // 3
        num = den = 0.0
        for _ in range(8):
            CL.run_until(dns, dns.time_stepping.t + 0.1)
            w, sg = CL.subgrid_terms(dns, les_oper)
            pred = folded.forward(w.float().reshape(1, nl, nl, 1))[0]
            num += float(((pred - sg.float()) ** 2).sum())
            den += float((sg.float() ** 2).sum())
    return {
        "workload": (f"ns2d forced (tcrandom k in [{band[0]}, {band[1]}], rate 1), nu_2 = {nu}; "
                     f"DNS {nd}^2 FP32; LES {nl}^2 FP32 with the 6-layer 5x5 32-channel "
                     f"network trained on {nruns} DNS runs x {nsnaps} filtered snapshots "
                     f"({train_steps} Adam steps, batch 32); evaluation over {t_eval} time "
                     "units from a held-out DNS"),
        "speedup_dns_over_les_closure": wall["dns"] / wall["les_closure"],
        "ms_per_step": {k: wall[k] / max(steps[k], 1) * 1e3 for k in wall},
        "steps": steps,
        "spectrum_error": {"shells_1_to_kmax": CL.spectrum_errors(spectra, kmax),
                           "small_scales": CL.spectrum_errors(spectra, kmax, lo=kmax // 2),
                           "kmax": kmax},
        "apriori_subgrid_nmse": num / den,
        "training_loss": [l1[0][1], l2[-1][1]],
        "timings_s": {"training_data": t_data, "training": t_train,
                      "total": time.perf_counter() - t00},
    }


def other_configs(args, torch, _lib, copy_gbps):
    """Short legs of the other workloads, in the default run so that every
    config is measured in one driver run: ns3d.strat 1024^3 (cfg 5, FP64
    oracle check at 128^3), the 512-run ensemble (cfg 3), the closure
    workload's 2048^2 DNS and the closure at reduced size (cfg 4), and the
    network's training step.  Each leg's failure is recorded, not raised."""
    import types

    legs = {}

    def leg(name, fn):
        print(f"[bench] leg {name}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        try:
            with contextlib.redirect_stdout(sys.stderr):
                legs[name] = fn()
        except Exception as e:  # noqa: BLE001 - recorded in the JSON line
            legs[name] = {"error": f"{type(e).__name__}: {e}"}
        legs[name]["leg_wall_s"] = time.perf_counter() - t0
        gc.collect()
        torch.cuda.empty_cache()

    def ns(**kw):
        base = dict(vars(args))
        base.update(kw)
        return types.SimpleNamespace(**base)

    leg("strat_1024", lambda: bench_strat(
        ns(n=1024, steps=10, warmup=2, precision="float32", cpu_n=128, oracle_steps=2,
           no_cpu_baseline=False), torch, _lib, copy_gbps))
    leg("ensemble_512_runs", lambda: bench_ensemble(
        ns(steps=100, warmup=3, no_cpu_baseline=True), torch, _lib, copy_gbps))
    leg("ns2d_2048_dns", lambda: bench_ns2d(
        ns(n=2048, steps=50, warmup=5, precision="float32"), torch, _lib, copy_gbps))
    leg("closure_reduced", lambda: closure_reduced(torch))
    leg("cnn_train", lambda: bench_cnn_train(ns(steps=20, warmup=3, cnn_batch=32,
                                                cnn_size=128), torch, _lib))
    return legs


def bench_cnn_train(args, torch, _lib):
    """Training of the closure network on the device: Adam steps of the
    6-layer 5x5 32-channel network at batch B on H x H images (synthetic
    inputs and targets), and the weight gradient of one hidden 32 -> 32
    layer (fsx_cnn_conv_wgrad: the MFMA kernel, the reduction of its
    partials and the bias gradient) timed alone with HIP events on the
    stream it runs on."""
    import ctypes

    from fluidsim_amd import closure as CL

    B, H = args.cnn_batch, args.cnn_size
    npix = B * H * H
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn((B, H, H, 1), generator=g).cuda()
    tgt = torch.randn((B, H, H), generator=g).cuda()
    net = CL.CNNClosure(n_in=1, nlayers=6, seed=0)
    for _ in range(args.warmup):
        net.train_step(x, tgt, want_loss=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.train_step(x, tgt, want_loss=False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # one hidden layer's weight gradient, alone
    a = torch.randint(-16000, 16000, (B, H, H, 32), dtype=torch.int16, generator=g).cuda()
    d = torch.randint(-16000, 16000, (B, H, H, 32), dtype=torch.int16, generator=g).cuda()
    nb = ctypes.c_longlong(0)
    _lib.call("fsx_cnn_wgrad_workspace", B, H, 32, ctypes.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    dw = torch.empty((32, 25, 32), dtype=torch.float32, device="cuda")
    db = torch.empty(32, dtype=torch.float32, device="cuda")
    s = _lib.stream_ptr()

    def wgrad():
        _lib.call("fsx_cnn_conv_wgrad", _lib.dptr(a), 32, _lib.dptr(d), 32, B, H, H,
                  _lib.dptr(dw), _lib.dptr(db), _lib.dptr(work), nb.value, s)

    for _ in range(3):
        wgrad()
    reps = 50
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        wgrad()
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / reps * 1e3
    flops = 2.0 * npix * 32 * 25 * 32
    tflops = flops / (us * 1e-6) / 1e12
    result = {
        "metric": "closure network training steps/s",
        "value": args.steps / elapsed,
        "unit": "Adam steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 x bf16 -> f32",
        "data": "synthetic",
        "config": {"workload": f"6-layer 5x5 32-channel GELU network, batch {B} of "
                               f"{H}x{H}, MSE loss, Adam", "parallelism": "single GPU"},
        "roofline": {
            "bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
            "frac": tflops / 2500.0, "traffic": None,
            "kernel": "fsx_cnn_conv_wgrad, hidden 32 -> 32 layer (k_conv_wgrad<32, 32> "
                      "+ reduction of the partials), HIP events",
            "us_per_call": us,
            "flops_per_call": flops,
        },
    }
    return result


def cpu_baseline_closure(net, nd, nl, nu, reps=2):
    """FP64 oracle on the host cores: one LES RK4 step with the network at
    nl^2 (the ns2d step plus 4 network forwards, numpy) and one DNS RK4 step
    at nd^2, each averaged over `reps`."""
    from oracle import cnn as C
    from oracle import operators as O
    from oracle.solvers import NS2D
    from fluidsim_amd.synthetic import ns2d_noise_rot_fft

    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    O.set_workers(cores)
    layers = []
    for l in range(net.nlayers):
        w = net.w[l].cpu().numpy().astype(np.float64)
        co = net.cout[l]
        layers.append((w[:co], net.b[l].cpu().numpy()[:co].astype(np.float64)))
    onet = C.Net(layers)

    def step_time(n):
        sim = NS2D(O.Oper2D(n, n), nu_2=nu, deltat0=1e-3)
        sim.init_from_rotfft(ns2d_noise_rot_fft(n, n, seed=0, kmax=12))
        sim.one_time_step_computation()
        t0 = time.perf_counter()
        for _ in range(reps):
            sim.one_time_step_computation()
        return (time.perf_counter() - t0) / reps

    t_les = step_time(nl)
    x = np.zeros((1, nl, nl, 8))
    x[0, :, :, 0] = np.random.default_rng(0).standard_normal((nl, nl))
    t0 = time.perf_counter()
    for _ in range(reps):
        onet.forward(x)
    t_net = (time.perf_counter() - t0) / reps
    t_dns = step_time(nd)
    t_les_net = t_les + 4 * t_net
    return {
        "value": 1.0 / t_les_net,
        "unit": f"LES RK4 steps/s ({nl}^2 with the network)",
        "cores": cores,
        "kind": "port",
        "sample": (f"FP64 oracle, {reps} steps each: ns2d {nl}^2 RK4 step {t_les * 1e3:.1f} ms "
                   f"+ 4 network forwards of {t_net * 1e3:.1f} ms (numpy) = "
                   f"{t_les_net * 1e3:.1f} ms per LES step; DNS {nd}^2 RK4 step "
                   f"{t_dns * 1e3:.1f} ms (scipy.fft workers={cores})"),
        "dns_steps_per_s": 1.0 / t_dns,
    }


def read_traffic(kernel_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None


T_START = time.perf_counter()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--precision", default="float32", choices=["float32", "float64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=256,
                    help="grid of the all-cores CPU baseline sample (one core: half)")
    ap.add_argument("--ensemble", action="store_true",
                    help="time the batched ns2d ensemble (512 runs) instead")
    ap.add_argument("--closure", action="store_true",
                    help="run the learned-closure LES workload instead")
    ap.add_argument("--closure-dns", type=int, default=2048)
    ap.add_argument("--closure-les", type=int, default=128)
    ap.add_argument("--closure-runs", type=int, default=8)
    ap.add_argument("--closure-snaps", type=int, default=100)
    ap.add_argument("--closure-snap-dt", type=float, default=0.1)
    ap.add_argument("--closure-spinup", type=float, default=3.0)
    ap.add_argument("--closure-train-steps", type=int, default=2000)
    ap.add_argument("--closure-eval-t", type=float, default=4.0)
    ap.add_argument("--closure-sample-dt", type=float, default=0.2)
    ap.add_argument("--cnn-train", action="store_true",
                    help="time the closure network's training step and weight gradient")
    ap.add_argument("--cnn-batch", type=int, default=32)
    ap.add_argument("--cnn-size", type=int, default=128)
    ap.add_argument("--ns2d", action="store_true",
                    help="time the forced ns2d DNS of the closure workload (use --n 2048)")
    ap.add_argument("--strat", action="store_true",
                    help="time the ns3d.strat forced workload (use --n 1024)")
    ap.add_argument("--oracle-steps", type=int, default=3,
                    help="steps of the FP64 oracle check of --strat")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP32-vs-FP64 accuracy runs of the ns3d bench")
    ap.add_argument("--accuracy-steps", type=int, default=100)
    ap.add_argument("--cfg2-n", type=int, default=512,
                    help="grid of the FP32 / FP64 legs of the ns3d workload")
    ap.add_argument("--no-cfg2", action="store_true",
                    help="skip the FP32 / FP64 legs at --cfg2-n")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the short legs of the other workloads")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MiB first "
                         "(calibration kernel for the HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(
            "bench.py measures one MI355X: this engine has no multi-GPU path in "
            "this build (slab/pencil sharding over RCCL is design-only). "
            "Run with --gpus 1.",
            file=sys.stderr,
        )
        sys.exit(2)

    import torch
    from fluidsim_amd import _lib

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n, prec = args.n, (4 if args.precision == "float32" else 8)
    copy_gbps = measured_copy_bandwidth(torch)
    mode = None
    if args.ensemble:
        mode = bench_ensemble(args, torch, _lib, copy_gbps)
    elif args.closure:
        mode = bench_closure(args, torch, _lib)
    elif args.cnn_train:
        mode = bench_cnn_train(args, torch, _lib)
    elif args.strat:
        mode = bench_strat(args, torch, _lib, copy_gbps)
    elif args.ns2d:
        mode = bench_ns2d(args, torch, _lib, copy_gbps)
    if mode is not None:
        print(json.dumps(mode), flush=True)
        return
    if args.copy_mb:
        src = torch.ones(args.copy_mb * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print(f"[bench] building ns3d {n}^3 {args.precision}", file=sys.stderr, flush=True)
    # the solver's own progress messages go to stderr: stdout is the JSON line
    t_build = time.perf_counter()
    with contextlib.redirect_stdout(sys.stderr):
        sim = make_sim(n, args.precision, device_init=n > 512)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    ts = sim.time_stepping
    plan = sim.oper.plan.handle
    print(f"[bench] dt = {ts.deltat:.6g}; warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ts.one_time_step_computation()
    torch.cuda.synchronize()

    _lib.call("fsx_timing_enable", plan, 1)
    ms8 = (_lib.ctypes.c_double * 8)()
    by8 = (_lib.ctypes.c_double * 8)()
    cn8 = (_lib.ctypes.c_longlong * 8)()
    _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts.one_time_step_computation()
        ts.t += ts.deltat
        ts.it += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("fsx_timing_read", plan, ms8, by8, cn8, 1)
    _lib.call("fsx_timing_enable", plan, 0)
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    steps_per_s = args.steps / elapsed
    diagnostics = time_diagnostics(sim, plan, torch)
    kernels = {}
    for k in range(7):
        if cn8[k]:
            kernels[KINDS[k]] = {
                "avg_ms": ms8[k] / cn8[k],
                "launches_per_step": cn8[k] / args.steps,
                "share_of_step": ms8[k] / (elapsed * 1e3),
                "alg_GBps": by8[k] / (ms8[k] * 1e-3) / 1e9,
                "alg_MB_per_launch": by8[k] / cn8[k] / 1e6,
            }
    dom = max(range(7), key=lambda k: ms8[k])
    achieved = by8[dom] / (ms8[dom] * 1e-3) / 1e9
    balg = b_alg_ns3d(n, prec)
    traffic = read_traffic(f"ns3d_{n}_{args.precision}:{KINDS[dom]}")

    result = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == 4 else "f64",
        "data": "synthetic",
        "config": {
            "workload": (
                f"ns3d {n}^3 {args.precision} RK4, nu_8 with k_max*eta=1, "
                "E(k)~k^4 exp(-2(k/4)^2) unit-energy solenoidal init (seed 0, "
                + ("device generator" if n > 512 else "host generator")
                + "), fixed dt from initial CFL 0.5, 2/3 dealiasing"
            ),
            "grid": [n, n, n],
            "parallelism": "single GPU",
            "setup_s": t_build,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": KINDS[dom],
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": traffic,
            "traffic_source": ("HBM bytes per launch of this kernel at this workload from "
                               "rocprofv3 FETCH_SIZE / WRITE_SIZE runs of bench.py, "
                               "committed in profiles/traffic.json (not measured in "
                               "this run)"),
            "copy_peak_measured": copy_gbps,
            "frac_of_copy": achieved / copy_gbps,
        },
        "step_roofline": {
            "B_alg_GB_per_step": balg / 1e9,
            "achieved_GBps": balg * steps_per_s / 1e9,
            "frac": balg * steps_per_s / 1e9 / PEAK_HBM_GBPS,
            "frac_of_copy": balg * steps_per_s / 1e9 / copy_gbps,
        },
        "kernels": kernels,
        "diagnostics": diagnostics,
    }
    del sim, ts
    gc.collect()
    torch.cuda.empty_cache()
    if not args.no_cfg2:
        m = args.cfg2_n
        print(f"[bench] {m}^3 FP32 and FP64 legs", file=sys.stderr, flush=True)
        result["cfg2"] = cfg2_legs(torch, m, steps_total=args.accuracy_steps)
    if not args.no_accuracy:
        print("[bench] accuracy: FP32 and FP64 engine vs the FP64 oracle", file=sys.stderr,
              flush=True)
        with contextlib.redirect_stdout(sys.stderr):
            result["accuracy"] = ns3d_accuracy(torch)
        if "cfg2" in result:
            result["accuracy"]["float32_vs_float64_engine"] = result["cfg2"].pop(
                "float32_vs_float64_engine")
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(n, args.cpu_n, args.cpu_n // 2)
    if not args.no_legs:
        result["other_configs"] = other_configs(args, torch, _lib, copy_gbps)
    result["bench_wall_s"] = time.perf_counter() - T_START
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
