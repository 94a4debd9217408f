"""Learned subgrid closure for coarse 2D LES (ns2d), on the matrix cores.

fluidsim has no learned closure: its LES hook is ``TurbModel.get_forcing``
(fluidsim/base/turb_model/base.py:87-88, Smagorinsky only, ns3d only); ns2d
calls no model (fluidsim/solvers/ns2d/solver.py:191-192).  This module adds

* ``CNNClosure``: 5x5 periodic convolutions (32 channels, GELU), BF16
  weights and activations with FP32 accumulation on MFMA (``fsx_cnn_*``),
  FP32 master weights trained with Adam on the device;
* ``subgrid_terms``: training pairs from a DNS state, filtered by the sharp
  spectral cut of fluidsim's resolution change (fill_field_fft_2d,
  fluidsim/base/init_fields.py:300-314): input the coarse vorticity, target
  the subgrid vorticity tendency S = cut(N(w)) - N(cut(w)) with N the ns2d
  nonlinear term (fluidsim/solvers/ns2d/solver.py:111-194);
* ``SimulLES``: the ns2d solver on the coarse grid with the network's S
  added to the nonlinear tendency at every Runge-Kutta stage.
"""

import ctypes

import numpy as np

from . import _lib


def _torch():
    import torch

    return torch


def bf16_bits(a):
    """Round-to-nearest-even BF16 bit patterns (int16 array) of float data."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r.view(np.int16)


def bf16_value(bits):
    return (np.asarray(bits).view(np.uint16).astype(np.uint32) << 16).view(np.float32)


class CNNClosure:
    """Convolutional network x (B, H, W, n_in) -> y (B, H, W) on the device.

    Layer l maps cin[l] (8 for the input layer, channel-padded, else 32)
    to cout[l] channels (32, the last layer 1) with a periodic 5x5
    convolution; GELU after every layer but the last.
    """

    def __init__(self, n_in=1, hidden=32, nlayers=6, seed=0, lr=1e-3):
        torch = _torch()
        if hidden != 32:
            raise ValueError("the MFMA kernels use 32 hidden channels")
        if not 1 <= n_in <= 8:
            raise ValueError("1 to 8 input channels")
        self.n_in = n_in
        self.nlayers = nlayers
        self.cin = [8] + [32] * (nlayers - 1)
        self.cout = [32] * (nlayers - 1) + [1]
        self.lr = lr
        rng = np.random.default_rng(seed)
        self.w, self.b = [], []
        for l in range(nlayers):
            fan_in = 25 * (n_in if l == 0 else 32)
            w = np.zeros((32, 25, self.cin[l]), np.float32)
            std = np.sqrt(2.0 / fan_in) * (0.1 if l == nlayers - 1 else 1.0)
            real_in = n_in if l == 0 else 32
            w[: self.cout[l], :, :real_in] = rng.standard_normal(
                (self.cout[l], 25, real_in)) * std
            self.w.append(torch.from_numpy(w).cuda())
            self.b.append(torch.zeros(32, dtype=torch.float32, device="cuda"))
        self.wb = [torch.empty(w.shape, dtype=torch.int16, device="cuda") for w in self.w]
        self._refresh_bf16()
        self.m1 = [torch.zeros_like(t) for t in self.w + self.b]
        self.m2 = [torch.zeros_like(t) for t in self.w + self.b]
        self.step_count = 0
        self._work = {}
        self.zero_bias = torch.zeros(32, dtype=torch.float32, device="cuda")

    # ------------------------------------------------------------ state
    def _refresh_bf16(self):
        for w, wb in zip(self.w, self.wb):
            _lib.call("fsx_cnn_to_bf16", _lib.dptr(w), w.numel(), _lib.dptr(wb),
                      _lib.stream_ptr())

    def state_dict(self):
        return {f"w{l}": self.w[l].cpu().numpy() for l in range(self.nlayers)} | {
            f"b{l}": self.b[l].cpu().numpy() for l in range(self.nlayers)}

    def load_state_dict(self, d):
        torch = _torch()
        for l in range(self.nlayers):
            self.w[l].copy_(torch.from_numpy(np.asarray(d[f"w{l}"], np.float32)))
            self.b[l].copy_(torch.from_numpy(np.asarray(d[f"b{l}"], np.float32)))
        self._refresh_bf16()

    # ------------------------------------------------------------ buffers
    def _buffers(self, B, H, W):
        torch = _torch()
        key = (B, H, W)
        if key not in self._work:
            npix = B * H * W
            self._work[key] = {
                "x": torch.empty((B, H, W, 8), dtype=torch.int16, device="cuda"),
                "fwd": torch.empty(2 * npix * 32 * 2, dtype=torch.uint8, device="cuda"),
            }
        return self._work[key]

    def pack_input(self, x):
        """(B, H, W, n_in) float -> BF16 (B, H, W, 8) channel-padded."""
        B, H, W, c = x.shape
        buf = self._buffers(B, H, W)["x"]
        xc = x.contiguous()  # held until the launch is issued
        _lib.call("fsx_cnn_act", _lib.dptr(xc), B * H * W, c, 8, 0, _lib.dptr(buf),
                  _lib.stream_ptr())
        return buf

    # ------------------------------------------------------------ inference
    def forward(self, x, out=None):
        """Network output (B, H, W) float32 for x (B, H, W, n_in) float32."""
        torch = _torch()
        B, H, W, _ = x.shape
        xb = self.pack_input(x)
        work = self._buffers(B, H, W)["fwd"]
        if out is None:
            out = torch.empty((B, H, W), dtype=torch.float32, device="cuda")
        nl, cin, cout, wp, bp = self.layer_args()
        _lib.call("fsx_cnn_forward", nl, cin, cout, wp, bp, _lib.dptr(xb), B, H, W,
                  _lib.dptr(out), _lib.dptr(work), work.numel(), _lib.stream_ptr())
        return out

    def layer_args(self):
        """(nl, cin, cout, BF16 weight pointers, bias pointers) as ctypes
        arrays; the tensors are updated in place, so the pointers hold."""
        if getattr(self, "_layer_args", None) is None:
            nl = self.nlayers
            self._layer_args = (
                nl,
                (ctypes.c_int * nl)(*self.cin),
                (ctypes.c_int * nl)(*self.cout),
                (ctypes.c_void_p * nl)(*[t.data_ptr() for t in self.wb]),
                (ctypes.c_void_p * nl)(*[t.data_ptr() for t in self.b]),
            )
        return self._layer_args

    # ------------------------------------------------------------ training
    def train_step(self, x, target, lr=None, want_loss=True):
        """One Adam step on the mean-squared error of x -> target."""
        torch = _torch()
        s = _lib.stream_ptr()
        B, H, W, _ = x.shape
        npix = B * H * W
        dev = dict(dtype=torch.float32, device="cuda")
        acts = [self.pack_input(x)]
        pres = []
        for l in range(self.nlayers):
            last = l == self.nlayers - 1
            y = torch.empty((B, H, W, self.cout[l]), **dev)
            if last:
                _lib.call("fsx_cnn_conv_fwd", _lib.dptr(acts[l]), B, H, W, self.cin[l],
                          _lib.dptr(self.wb[l]), _lib.dptr(self.b[l]), self.cout[l],
                          _lib.dptr(y), s)
            else:
                # pre-activations (for the backward pass) and the BF16 GELU
                # activations from one launch
                a = torch.empty((B, H, W, 32), dtype=torch.int16, device="cuda")
                _lib.call("fsx_cnn_conv_fwd_act", _lib.dptr(acts[l]), B, H, W, self.cin[l],
                          _lib.dptr(self.wb[l]), _lib.dptr(self.b[l]), _lib.dptr(y),
                          _lib.dptr(a), s)
                acts.append(a)
            pres.append(y)
        # loss gradient w.r.t. the output, BF16 with 8 channels
        d = torch.empty((B, H, W, 8), dtype=torch.int16, device="cuda")
        part = torch.empty(256, dtype=torch.float64, device="cuda")
        loss = ctypes.c_double(0.0)
        tc = target.contiguous()
        _lib.call("fsx_cnn_mse", _lib.dptr(pres[-1]), _lib.dptr(tc), npix, 8,
                  _lib.dptr(d), _lib.dptr(part), ctypes.byref(loss) if want_loss else None, s)
        grads_w = [None] * self.nlayers
        grads_b = [None] * self.nlayers
        nbytes = ctypes.c_longlong(0)
        for l in range(self.nlayers - 1, -1, -1):
            ci, co_pad = self.cin[l], (8 if l == self.nlayers - 1 else 32)
            _lib.call("fsx_cnn_wgrad_workspace", B, H, ci, ctypes.byref(nbytes))
            work = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
            gw = torch.empty((32, 25, ci), **dev)
            gb = torch.empty(32, **dev)
            _lib.call("fsx_cnn_conv_wgrad", _lib.dptr(acts[l]), ci, _lib.dptr(d), co_pad, B, H,
                      W, _lib.dptr(gw), _lib.dptr(gb), _lib.dptr(work), nbytes.value, s)
            grads_w[l], grads_b[l] = gw, gb
            if l == 0:
                break
            # input gradient: flipped, transposed weights on d
            wt = torch.empty((32, 25, co_pad), dtype=torch.int16, device="cuda")
            _lib.call("fsx_cnn_conv_flip", _lib.dptr(self.wb[l]), ci, self.cout[l], co_pad,
                      _lib.dptr(wt), s)
            # and through the previous layer's GELU, in the convolution epilogue
            dn = torch.empty((B, H, W, 32), dtype=torch.int16, device="cuda")
            _lib.call("fsx_cnn_conv_bwd_act", _lib.dptr(d), B, H, W, co_pad, _lib.dptr(wt),
                      _lib.dptr(self.zero_bias), _lib.dptr(pres[l - 1]), _lib.dptr(dn), s)
            d = dn
        self.step_count += 1
        lr = self.lr if lr is None else lr
        # the weight and bias tensors, up to 16 per launch
        params = self.w + self.b
        grads = grads_w + grads_b
        wbs = self.wb + [None] * self.nlayers
        for c0 in range(0, len(params), 16):
            sl = slice(c0, c0 + 16)
            nt = len(params[sl])

            def ptrs(ts):
                return (ctypes.c_void_p * nt)(
                    *[t.data_ptr() if t is not None else None for t in ts[sl]])

            _lib.call("fsx_cnn_adam_multi", nt, ptrs(params), ptrs(grads), ptrs(self.m1),
                      ptrs(self.m2), (ctypes.c_int * nt)(*[p.numel() for p in params[sl]]),
                      ptrs(wbs), float(lr), 0.9, 0.999, 1e-8, self.step_count, s)
        self.last_grads = (grads_w, grads_b)
        return loss.value if want_loss else None

    # ------------------------------------------------------------ LES use
    def folded(self, in_scale, out_scale):
        """Copy whose input is the raw field and output the raw subgrid term:
        y = out_scale * net(x / in_scale) (the scales enter the first and
        last layers, which are linear in their inputs / outputs)."""
        torch = _torch()
        other = CNNClosure.__new__(CNNClosure)
        other.__dict__.update({k: v for k, v in self.__dict__.items()
                               if k not in ("w", "b", "wb", "_work", "m1", "m2", "_layer_args")})
        other.w = [t.clone() for t in self.w]
        other.b = [t.clone() for t in self.b]
        other.w[0] /= float(in_scale)
        other.w[-1] *= float(out_scale)
        other.b[-1] *= float(out_scale)
        other.wb = [torch.empty(w.shape, dtype=torch.int16, device="cuda") for w in other.w]
        other._work = {}
        other.m1 = other.m2 = []
        other._refresh_bf16()
        return other


# ---------------------------------------------------------------- data
def _resize(src, src_shapeK, dst, dst_shapeK, prec):
    _lib.call("fsx_resize_fft", _lib.dptr(src), 1, src_shapeK[0], src_shapeK[1], _lib.dptr(dst),
              1, dst_shapeK[0], dst_shapeK[1], 1, prec, _lib.stream_ptr())


def cut_to_les(dns_oper, les_oper, a_fft):
    """Sharp spectral cut of a DNS half spectrum to the LES grid
    (fill_field_fft_2d, fluidsim/base/init_fields.py:300-314), then the LES
    grid's own dealiasing: the filter whose output the LES state is."""
    torch = _torch()
    out = torch.zeros(tuple(les_oper.shapeK), dtype=a_fft.dtype, device="cuda")
    a = a_fft.reshape(tuple(dns_oper.shapeK)).contiguous()
    _resize(a, dns_oper.shapeK, out, les_oper.shapeK, dns_oper.precision)
    les_oper.dealiasing(out)
    return out


def _combine(oper, out, terms):
    """out = sum_j a_j x_j over one field (fsx_if_combine with b = e = p = 0)."""
    n = len(terms)
    ins = (ctypes.c_void_p * n)(*[t[0].data_ptr() for t in terms])
    a = _lib.doubles([float(t[1]) for t in terms])
    z = _lib.doubles([0.0] * n)
    pw = (ctypes.c_int * n)(*([0] * n))
    _lib.call("fsx_if_combine", oper.plan.handle, 1, _lib.dptr(out), n, ins, a, z, z, pw,
              _lib.doubles([0.0, 0.0, 0.0]), _lib.doubles([0.0] * 4), 0, _lib.stream_ptr())


def subgrid_terms(dns_sim, les_oper, beta=0.0):
    """(w_bar, S) on the LES grid (real fields) from the DNS state:
    w_bar = cut(w), S = cut(N(w)) - N(w_bar) with N the dealiased ns2d
    nonlinear term of each grid."""
    torch = _torch()
    od = dns_sim.oper
    rot = dns_sim.state.state_spect.data
    N = torch.zeros_like(rot)
    od.ns2d_tendency(rot, N, beta)
    wk = cut_to_les(od, les_oper, rot[0])
    Nk = cut_to_les(od, les_oper, N[0])
    Nl = torch.zeros((1,) + tuple(les_oper.shapeK), dtype=rot.dtype, device="cuda")
    wk1 = wk.reshape(Nl.shape)
    les_oper.ns2d_tendency(wk1, Nl, beta)
    S = torch.empty_like(Nk)
    _combine(les_oper, S, [(Nk, 1.0), (Nl[0], -1.0)])
    return les_oper.ifft(wk), les_oper.ifft(S)


# ---------------------------------------------------------------- LES
def make_les_class():
    """The ns2d solver with the learned subgrid term added to the
    nonlinear tendency of every Runge-Kutta stage (fsx_ns2d_rk_stage_ext)."""
    from .solvers.ns2d.solver import Simul as SimulNS2D

    class SimulLES(SimulNS2D):
        closure_net = None  # CNNClosure with folded scales (raw w -> raw S)

        def _graph_key(self):
            """A captured step holds the network's weight and work buffers:
            one graph per network object."""
            return id(self.closure_net) if self.closure_net is not None else None

        def _closure_term(self, src):
            """Network term in spectral space from the vorticity src[0]: one
            engine call (c2r, network, r2c, dealiasing)."""
            torch = _torch()
            net = self.closure_net
            if net.n_in != 1:
                raise ValueError("the LES closure network takes the vorticity only")
            plan = self.oper.plan.handle
            if getattr(self, "_cl_k", None) is None:
                nb = ctypes.c_longlong(0)
                _lib.call("fsx_ns2d_closure_workspace", plan, nb)
                self._cl_work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
                self._cl_k = torch.empty((1,) + tuple(self.oper.shapeK), dtype=src.dtype,
                                         device="cuda")
            nl, cin, cout, wp, bp = net.layer_args()
            _lib.call("fsx_ns2d_closure_term", plan, nl, cin, cout, wp, bp, _lib.dptr(src),
                      _lib.dptr(self._cl_k), _lib.dptr(self._cl_work), self._cl_work.numel(),
                      _lib.stream_ptr())
            return self._cl_k

        def _engine_rk_stage(self, scheme, stage, S, tmp, t1, visc, cfl, dt_max):
            if self.closure_net is None:
                return super()._engine_rk_stage(scheme, stage, S, tmp, t1, visc, cfl, dt_max)
            ext = self._closure_term(S if stage == 0 else t1)
            _lib.call("fsx_ns2d_rk_stage_ext", self.oper.plan.handle, scheme, stage,
                      _lib.dptr(S), _lib.dptr(tmp), _lib.dptr(t1), float(self.params.beta),
                      visc, cfl, dt_max, _lib.dptr(ext), _lib.stream_ptr())

    return SimulLES


def ns2d_params(n, nu_2, precision="float32", forcing=True, nkmin=8, nkmax=10, rate=1.0,
                les=False, graph=False):
    """Parameters of the forced ns2d runs of the closure workload (graph:
    steps replayed from a HIP graph, params.time_stepping.hip_graph)."""
    from .solvers.ns2d.solver import Simul

    cls = make_les_class() if les else Simul
    params = cls.create_default_params()
    params.oper.nx = params.oper.ny = n
    params.oper.Lx = params.oper.Ly = 2 * np.pi
    params.oper.precision = precision
    params.nu_2 = nu_2
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.time_stepping.hip_graph = graph
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    if forcing:
        params.forcing.enable = True
        params.forcing.type = "tcrandom"
        params.forcing.nkmin_forcing = nkmin
        params.forcing.nkmax_forcing = nkmax
        params.forcing.forcing_rate = rate
    return cls, params


def share_forcing(src, sims):
    """Give the runs `sims` the forcing of `src`: the coarse forcing box does
    not depend on the grid, so each maker takes src's random pair and renewal
    time (fluidsim/base/forcing/specific.py:768-868); every run then draws
    its renewals from its own copy of numpy's random state."""
    mk0 = src.forcing.forcing_maker
    for sim in sims:
        mk = sim.forcing.forcing_maker
        if mk.forcing0.shape != mk0.forcing0.shape:
            raise ValueError("the runs' coarse forcing boxes differ")
        mk.t_last_change = mk0.t_last_change
        mk._seed0, mk._seed1 = mk0._seed0, mk0._seed1
        mk.forcing0, mk.forcing1 = mk0.forcing0.copy(), mk0.forcing1.copy()
        mk._upload_pair(mk.forcing0, mk.forcing1)


def run_until(sim, t):
    """Advance a Simul until its time reaches t; returns the steps taken."""
    ts = sim.time_stepping
    if not getattr(ts, "_prepare_main_loop_called", False):
        ts.prepare_main_loop()
    n = 0
    while ts.t < t - 1e-12:
        ts.one_time_step()
        n += 1
    return n


def compare_runs(runs, les_oper, t_eval, sample_dt, dns_key="dns"):
    """Advance the runs (name -> Simul, started at the same time) over t_eval
    and sample their energy spectra on the LES grid every sample_dt (the
    DNS cut to the LES grid); each run draws its forcing renewals from its
    own copy of numpy's random state.  Returns the sampled spectra, the wall
    time and the steps of each run."""
    torch = _torch()
    rng_state = {k: np.random.get_state() for k in runs}
    spectra = {k: [] for k in runs}
    wall = {k: 0.0 for k in runs}
    steps = {k: 0 for k in runs}
    import time as _time

    t_start = next(iter(runs.values())).time_stepping.t
    nsamp = int(round(t_eval / sample_dt))
    for j in range(nsamp):
        t_next = t_start + (j + 1) * sample_dt
        for k, sim in runs.items():
            np.random.set_state(rng_state[k])
            torch.cuda.synchronize()
            t0 = _time.perf_counter()
            steps[k] += run_until(sim, t_next)
            torch.cuda.synchronize()
            wall[k] += _time.perf_counter() - t0
            rng_state[k] = np.random.get_state()
            rot = sim.state.state_spect.data
            if k == dns_key:
                rot = cut_to_les(sim.oper, les_oper, rot[0])
            spectra[k].append(energy_spectrum(les_oper, rot))
        for k, sim in runs.items():
            if not np.isfinite(spectra[k][-1]).all():
                raise ValueError(f"{k}: non-finite spectrum at t = {sim.time_stepping.t}")
    return spectra, wall, steps


def spectrum_errors(spectra, kmax, lo=0, dns_key="dns"):
    """Relative L2 error of each run's time-averaged energy spectrum against
    the DNS's over shells 1 + lo .. kmax."""
    mean = {k: np.mean(v, axis=0)[1: kmax + 1] for k, v in spectra.items()}
    ref = mean[dns_key][lo:]
    return {k: float(np.linalg.norm(m[lo:] - ref) / np.linalg.norm(ref))
            for k, m in mean.items() if k != dns_key}


def run_steps(sim, nsteps):
    """Advance a Simul by nsteps (the time loop of TimeStepping.main_loop)."""
    ts = sim.time_stepping
    if not getattr(ts, "_prepare_main_loop_called", False):
        ts.prepare_main_loop()
    for _ in range(nsteps):
        ts.one_time_step()


def energy_spectrum(oper, rot_fft):
    """2D (shell) energy spectrum of a vorticity field on the device (FP64)."""
    from .base import diagnostics as dg

    r = rot_fft.reshape((1,) + tuple(oper.shapeK)).contiguous()
    return dg.run(oper, dg.NS2D_SPECTRA, dg.SHELL, [r[0]])["shell"][0]


def train(net, samples_w, samples_s, steps, batch=16, lr=1e-3, seed=0, log=None,
          symmetries=True):
    """Adam on (w / sigma_w) -> (S / sigma_S) with random periodic shifts;
    returns (sigma_w, sigma_S, losses).

    With ``symmetries``, each batch is also turned by a random multiple of
    90 degrees and mirrored at random: the 2D Navier-Stokes tendency commutes
    with the rotations of a square periodic box, and a mirror flips the sign
    of the vorticity and of its tendency alike (w is a pseudoscalar)."""
    torch = _torch()
    W = torch.stack(samples_w).float()
    Sx = torch.stack(samples_s).float()
    sw = float(W.std())
    ss = float(Sx.std())
    W /= sw
    Sx /= ss
    g = torch.Generator(device="cpu").manual_seed(seed)
    n, H, Wd = W.shape
    losses = []
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g)
        dy, dx = (int(v) for v in torch.randint(0, H, (2,), generator=g))
        xb = torch.roll(W[idx], (dy, dx), dims=(1, 2))
        yb = torch.roll(Sx[idx], (dy, dx), dims=(1, 2))
        if symmetries and H == Wd:
            turn, mirror = (int(v) for v in torch.randint(0, 4, (2,), generator=g))
            if turn:
                xb, yb = torch.rot90(xb, turn, (1, 2)), torch.rot90(yb, turn, (1, 2))
            if mirror % 2:
                xb, yb = -torch.flip(xb, (2,)), -torch.flip(yb, (2,))
        xb = xb.unsqueeze(-1).contiguous()
        yb = yb.contiguous()
        want = log is not None and (it % log == 0 or it == steps - 1)
        loss = net.train_step(xb, yb, lr=lr, want_loss=want)
        if want:
            losses.append((it, loss))
    return sw, ss, losses
