/* C-ABI of the fluidsim_amd pseudo-spectral engine (MI355X / gfx950).
 *
 * One plan per operator (grid, box, precision, dealiasing mask).  Device
 * buffers are plain pointers to HBM allocated by the caller (the Python host
 * passes torch tensors' data_ptr()); complex arrays are interleaved
 * (re, im) of the plan precision, in the sequential fluidfft layout:
 *   2D  X (ny, nx) real       <->  K (ny, nx/2+1) complex
 *   3D  X (nz, ny, nx) real   <->  K (nz, ny, nx/2+1) complex
 * Several fields are contiguous, one array after the other.
 * Streams are hipStream_t passed as void* (NULL = default stream).
 * Every function returns 0 on success or a negative FSX_ERR_* code;
 * fsx_last_error() copies the message of the last failure of the thread.
 * Calls are stream-ordered; only the entry points marked (sync), such as
 * fsx_get_scalars, fsx_sums, fsx_diag and fsx_set_dt, synchronise the
 * stream.  A plan must not be used from two host threads at once.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the fluidsim source tree).
 */
#ifndef FLUIDSIM_AMD_H
#define FLUIDSIM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FSX_ABI_VERSION 1

#define FSX_OK 0
#define FSX_ERR_ARG (-1)
#define FSX_ERR_HIP (-2)

typedef struct fsx_plan_s* fsx_plan;
typedef struct fsx_graph_s* fsx_graph;

/* ABI version (FSX_ABI_VERSION). */
int fsx_version(void);

/* Message of the last failure on this thread (NUL-terminated, truncated). */
int fsx_last_error(char* buf, int len);

/* Create an operator plan.
 * Replaces the construction of OperatorsPseudoSpectral2D / 3D
 * (fluidsim/operators/operators2d.py:115-193,
 *  fluidsim/operators/operators3d.py:207-278): grid, box, FFT plans and the
 * dealiasing mask.  where_dealiased is the host uint8 mask of shape K
 * (fluidfft's where_dealiased, possibly modified by truncation_shape,
 * NO_SHEAR_MODES, NO_KY0) or NULL for no dealiasing.  precision is 4 or 8
 * (bytes per real).  nf_max is the number of fields the workspace holds
 * (4 for ns2d, 6 for ns3d, 7 for ns3d.strat).  Any grid size >= 2 per axis:
 * power-of-two lengths run the register/LDS Stockham passes, other lengths
 * (as FFTW takes them) mixed-radix passes (radix 8, 4, 2, 3, 5 and a direct
 * DFT for other primes).  All device memory of the plan is allocated here,
 * including the buffers of fsx_diag. */
int fsx_plan_create(fsx_plan* out, int dim, int nx, int ny, int nz, double Lx,
                    double Ly, double Lz, int precision,
                    const unsigned char* where_dealiased, int nf_max);
int fsx_plan_destroy(fsx_plan p);

/* info[0..11] = dim, precision, nx, ny, nz, nkx, kept kx count, kept ky
 * (lo, hi), kept kz (lo, hi), mask-mode flag. */
int fsx_plan_info(fsx_plan p, int* info, int n);

/* Forward real transform divided by the number of points, nfields fields.
 * Replaces fft2 / fft3d / fft_as_arg
 * (fluidfft, called at fluidsim/solvers/ns2d/solver.py:181,
 *  fluidsim/solvers/ns3d/solver.py:237-239). */
int fsx_fft(fsx_plan p, const void* x, void* x_fft, int nfields, void* stream);

/* Unnormalised inverse real transform; x_fft is not modified.  The
 * imaginary parts of the kx = 0 and kx = nx/2 columns are ignored after the
 * y/z passes (numpy/pocketfft c2r semantics).
 * Replaces ifft2 / ifft3d / ifft_as_arg / ifft_as_arg_destroy
 * (fluidsim/solvers/ns2d/solver.py:161-171,
 *  fluidsim/solvers/ns3d/solver.py:210-220). */
int fsx_ifft(fsx_plan p, const void* x_fft, void* x, int nfields, void* stream);

/* ns2d nonlinear tendency N = dealias(fft(-u d_x rot - v (d_y rot + beta))).
 * Replaces Simul.tendencies_nonlin + compute_Frot
 * (fluidsim/solvers/ns2d/solver.py:33-39, 111-194).  Only non-dealiased
 * modes of out_fft are written (dealiased ones inside the kept box as 0);
 * the caller keeps out_fft zero outside the box (fsx_dealias with
 * box_only = 1).  want_maxima stores max|ux|, max|uy| for fsx_cfl_update. */
int fsx_ns2d_tendency(fsx_plan p, const void* rot_fft, void* out_fft,
                      double beta, int want_maxima, void* stream);

/* ns3d nonlinear tendency fft(v x omega) with omega = i k x v (+ f at k = 0),
 * dealiased, NOT projected (fsx_project or the projection folded into
 * fsx_rk_stage does it).  v_fft and out_fft hold 3 fields.
 * Replaces Simul.tendencies_nonlin (fluidsim/solvers/ns3d/solver.py:180-253)
 * up to project_state_spect. */
int fsx_ns3d_tendency(fsx_plan p, const void* v_fft, void* out_fft, int has_f,
                      double f, int want_maxima, void* stream);

/* One integrating-factor Runge-Kutta stage over the kept modes, with
 * E = exp(-dt sigma(k)), E2 = exp(-dt/2 sigma(k)) computed in FP64 from the
 * device time step (fsx_set_dt / fsx_cfl_update), sigma from
 * visc = {nu_2, nu_4, nu_8, nu_m4}.  scheme 4 (stages 0..3) or 2 (0..1).
 * project (3D): 0 none, 1 Leray (project_perpk3d), 2 toroidal, 3 poloidal,
 * applied to N and, at the last stage, to the new state.
 * Replaces _time_step_RK4 / _time_step_RK2 with ExactLinearCoefs
 * (fluidsim/base/time_stepping/pseudo_spect.py:103-152, 466-517, 798-984)
 * and the projection of one_time_step_computation
 * (fluidsim/solvers/ns3d/time_stepping.py:8-20).  Sets the NaN flag read by
 * fsx_get_scalars (pseudo_spect.py:241-243). */
int fsx_rk_stage(fsx_plan p, int scheme, int stage, int nvar, void* S,
                 void* tmp, void* t1, void* N, int project, int no_vz_kz0,
                 const double* visc, void* stream);

/* Fused Runge-Kutta stage of ns2d: N(src) with src = S at stage 0 and t1
 * otherwise, then the stage algebra of fsx_rk_stage applied in the last
 * forward pass, so the tendency never goes to HBM.  cfl > 0 at stage 0: the
 * maxima of |u|, |v| of this tendency update the device dt first
 * (fsx_cfl_update rule, deltat_max = dt_max).
 * Replaces one stage of _time_step_RK4 / _time_step_RK2 with
 * Simul.tendencies_nonlin (fluidsim/base/time_stepping/pseudo_spect.py
 * :466-517, 798-984; fluidsim/solvers/ns2d/solver.py:111-194) and the CFL of
 * fluidsim/base/time_stepping/base.py:320-366. */
int fsx_ns2d_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                      void* t1, double beta, const double* visc, double cfl,
                      double dt_max, void* stream);

/* fsx_ns2d_rk_stage with a full-size additive tendency ext_fft (one K
 * field of the plan, NULL: none) added to N before the stage algebra, like
 * the forcing (fluidsim/solvers/ns2d/solver.py:190-191): the hook of the
 * learned LES closure (no counterpart in the reference). */
int fsx_ns2d_rk_stage_ext(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                          void* t1, double beta, const double* visc, double cfl,
                          double dt_max, const void* ext_fft, void* stream);

/* Fused Runge-Kutta stage of ns3d (3 velocity fields), same contract as
 * fsx_ns2d_rk_stage with the tendency of fluidsim/solvers/ns3d/solver.py
 * :180-253 and the projection of fsx_rk_stage. */
int fsx_ns3d_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                      void* t1, int has_f, double f, int project, int no_vz_kz0,
                      const double* visc, double cfl, double dt_max,
                      void* stream);

/* Fused Runge-Kutta stage of ns3d.strat (4 fields: vx, vy, vz, b) with
 * Brunt-Vaisala frequency N: tendency
 *   N_v = P(fft(v x omega) + b e_z),  N_b = -i k.fft(v b) - N^2 v_z
 * (fluidsim/solvers/ns3d/strat/solver.py:29-33, 138-216), same contract as
 * fsx_ns3d_rk_stage. */
int fsx_strat_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                       void* t1, int has_f, double f, double N, int project,
                       int no_vz_kz0, const double* visc, double cfl,
                       double dt_max, void* stream);

/* ns3d.strat tendency of `state` (4 fields) into out_fft (4 fields),
 * projected (kind `project`) and dealiased; only kept modes are written.
 * Replaces Simul.tendencies_nonlin of
 * fluidsim/solvers/ns3d/strat/solver.py:138-216.  want_maxima stores
 * max|v_i| for fsx_cfl_update. */
int fsx_strat_tendency(fsx_plan p, const void* state, void* out_fft, int has_f,
                       double f, double N, int project, int no_vz_kz0,
                       int want_maxima, void* stream);

/* In-place projection of (a, b, c) over all modes: kind 1 project_perpk3d,
 * 2 project_toroidal, 3 project_poloidal
 * (fluidsim/operators/operators3d.py:787-937, fluidfft project_perpk3d). */
int fsx_project(fsx_plan p, void* a, void* b, void* c, int kind, void* stream);

/* Zero dealiased modes of nfields fields (box_only: only outside the
 * bounding box of kept modes).
 * Replaces dealiasing (fluidsim/operators/operators2d.py:200-220,
 * fluidsim/operators/operators3d.py:38-71, 336-342). */
int fsx_dealias(fsx_plan p, void* arr, int nfields, int box_only, void* stream);

/* Device time step used by fsx_rk_stage (synchronous). */
int fsx_set_dt(fsx_plan p, double dt, void* stream);

/* Zero the CFL maxima slots before a tendency with want_maxima. */
int fsx_reset_maxima(fsx_plan p, void* stream);

/* dt update from the maxima on the device
 * (fluidsim/base/time_stepping/base.py:320-366): ncomp velocity components,
 * CFL coefficient, deltat_max, 2 % hysteresis. */
int fsx_cfl_update(fsx_plan p, int ncomp, double cfl, double dt_max,
                   void* stream);

/* host6 = {max|u_0|, max|u_1|, max|u_2|, (unused), dt, nan flag} (sync). */
int fsx_get_scalars(fsx_plan p, double* host6, void* stream);
int fsx_clear_nan_flag(fsx_plan p, void* stream);

/* FP64 Parseval-weighted sums (sync), host_out6:
 * kind 0 (ns2d, f0 = rot): E, Z, epsK, epsK_hypo, epsZ, epsZ_hypo
 *   (fluidsim/solvers/ns2d/output/spatial_means.py:23-38);
 * kind 1 (f0..f2 = v): E, epsK, epsK_hypo
 *   (fluidsim/solvers/ns3d/output/spatial_means.py:23-73);
 * kind 2 (f0..f2 = v, f3 = b, N2 = N^2): E_K, E_A, epsK, epsA. */
int fsx_sums(fsx_plan p, int kind, const void* f0, const void* f1,
             const void* f2, const void* f3, double N2, const double* visc,
             double* host_out6, void* stream);

/* ---- forcing on a coarse spectral box ----------------------------------
 * fluidsim computes its forcing on a coarse grid and adds the full-size
 * copy to the tendencies at every stage
 * (fluidsim/base/forcing/specific.py:135-328, put_forcingc_in_forcing
 * 279-290; fluidsim/solvers/ns2d/solver.py:190-191,
 * fluidsim/solvers/ns3d/solver.py:243-244).  The engine keeps the coarse
 * box Fc (nvar fields of nzc x nyc x nkxc complex values of the plan
 * precision, nzc = 1 in 2D) on the device; while enabled, every Runge-Kutta
 * stage adds Fc to the tendency of the modes inside the box (before the
 * projection), the tendency-only entry points do not.
 *
 * fsx_forcing_setup: box shape, the forced modes (linear indices into the
 * box) with their Parseval weights, and the forced key: key_kind 0 = state
 * variable key_var, 1 = toroidal velocity vt_fft (3D; vecfft_from_vtfft,
 * fluidsim/operators/operators3d.py:961-992).  Fc starts at zero. */
int fsx_forcing_setup(fsx_plan p, int nyc, int nzc, int nkxc, int nvar,
                      int nmodes, const int* modes, const double* weights,
                      int key_kind, int key_var);

/* Raw random forcings f0, f1 of the key at the forced modes (host arrays of
 * nmodes interleaved complex128), TimeCorrelatedRandomPseudoSpectral's
 * forcing0 / forcing1 (fluidsim/base/forcing/specific.py:768-868).
 * Synchronises the stream first. */
int fsx_forcing_set_pair(fsx_plan p, const double* f0, const double* f1,
                         void* stream);

/* Whole coarse box from the host (stationary forcings; plan precision).
 * Synchronises the stream first. */
int fsx_forcing_set_coarse(fsx_plan p, const void* fc, void* stream);

/* Whether the Runge-Kutta stages add Fc. */
int fsx_forcing_enable(fsx_plan p, int on);

/* Normalisation of the next step (NormalizedForcing.compute,
 * fluidsim/base/forcing/specific.py:460-500, 587-726): with
 * f_r = f1 - coef (f1 - f0), a = dt/2 sum_w |f_r|^2 and
 * b = sum_w Re(conj(v) f_r) over the forced modes (v: the forced key of the
 * state, dt: the device time step), Fc = R f_r with R the root of
 * a R^2 + b R - forcing_rate = 0 chosen by which_root (0 minabs, 1 first,
 * 2 second, 3 positive).  fsx_forcing_step records the parameters; the
 * next stage-0 call of fsx_ns2d/ns3d/strat_rk_stage normalises on S after
 * its CFL update.  fsx_forcing_normalize does it at once on S (unfused
 * stepping: after fsx_cfl_update, before fsx_rk_stage). */
int fsx_forcing_step(fsx_plan p, double coef, double forcing_rate, int which_root);
int fsx_forcing_normalize(fsx_plan p, const void* S, void* stream);

/* Graph mode of the normalisation (no reference counterpart: support for
 * steps captured in a HIP graph, whose kernel arguments are fixed at
 * capture).  fsx_forcing_graph(p, 1): the normalisation reads coef,
 * forcing_rate and which_root from device memory; fsx_forcing_push writes
 * the values of the last fsx_forcing_step there, stream-ordered (call it on
 * the step's stream before each replay). */
int fsx_forcing_graph(fsx_plan p, int on);
int fsx_forcing_push(fsx_plan p, void* stream);

/* out[v] += scale * Fc[v] for v < nvar, at the full-grid modes of the box
 * (put_coarse_array_in_array_fft, fluidsim/operators/operators2d.py
 * :525-600): the forcing added to tendencies_nonlin, or, on a zeroed array,
 * get_forcing(). */
int fsx_forcing_add(fsx_plan p, void* out_fft, int nvar, double scale, void* stream);

/* ---- Smagorinsky turbulence model (3D plans) ----------------------------
 * SmagorinskyModel.get_forcing (fluidsim/base/turb_model/smagorinsky.py
 * :52-110, stress_tensor.py:15-44), in three pointwise steps around the
 * plan's transforms; fields ordered (xx, yy, zz, yx, zx, zy):
 *   fsx_strain_rate:      S_fft[6] = i (k_i v_j + k_j v_i) / 2 from v_fft[3];
 *   fsx_ifft(S_fft -> S, 6 fields);
 *   fsx_smagorinsky_phys: S[6] <- coef |S| S_ij in place, |S| = sqrt(S_ij S_ij)
 *                         (coef = 2 C_nuT, C_nuT = C (Lx/nx)^2 sqrt(2));
 *   fsx_fft(S -> g_fft, 6 fields);
 *   fsx_smagorinsky_div:  out[3] (+)= 2i sum_j k_j g_fft_ij
 *                         (accumulate != 0 adds to out). */
int fsx_strain_rate(fsx_plan p, const void* v_fft, void* S_fft, void* stream);
int fsx_smagorinsky_phys(fsx_plan p, void* S, double coef, void* stream);
int fsx_smagorinsky_div(fsx_plan p, const void* g_fft, void* out_fft, int accumulate,
                        void* stream);

/* Integrating-factor linear combination over all modes of nvar fields:
 *   out = sum_{j<nin} (a_j + b_j dt) exp(-e_j dt sigma(k))
 *                     exp(i phase_pow_j (shift . k)) in_j
 * (dt: the device time step, sigma from visc = {nu2, nu4, nu8, nu_m4}),
 * the step of the Euler, RK2-trapezoid and phase-shifted schemes
 * (step_Euler, step_like_RK2, mean_with_phaseshift, mul, div_inplace;
 * fluidsim/base/time_stepping/pseudo_spect.py:47-97, 245-796).  nin <= 4,
 * phase_pow in {-1, 0, 1}; out may alias an input; check_nan sets the NaN
 * flag read by fsx_get_scalars when field 0 of out holds a NaN. */
int fsx_if_combine(fsx_plan p, int nvar, void* out, int nin, void* const* ins,
                   const double* a, const double* b, const double* e,
                   const int* phase_pow, const double* shift, const double* visc,
                   int check_nan, void* stream);

/* Resolution change of nfields K arrays (no plan needed; nz = 1 in 2D):
 * fill_field_fft_2d (fluidsim/base/init_fields.py:300-314) and
 * fill_field_fft_3d (fluidsim/util/mini_oper_modif_resol.py:26-53).  Output
 * modes whose signed wavenumbers satisfy -m/2 < k <= m/2 (m = the smaller
 * size) on ky, kz and kx < min(nkx) take the input value; the others are
 * left unchanged (zero the output first for a truncation / padding).
 * precision 4 or 8 (complex64 / complex128). */
int fsx_resize_fft(const void* in, int nz_in, int ny_in, int nkx_in, void* out,
                   int nz_out, int ny_out, int nkx_out, int nfields, int precision,
                   void* stream);

/* ---- diagnostics: spectra, spectral energy budgets, spatial means ------
 * One pass over the K arrays with FP64 accumulation.  kind selects the
 * per-mode quantities (Q of them) and the input fields:
 *   0 ns2d spectra     (rot)                      Q=1  E
 *     fluidsim/solvers/ns2d/output/spectra.py:16-28
 *   1 ns2d budget      (rot, Frot, Fx, Fy)        Q=2  transferE, transferZ
 *     fluidsim/solvers/ns2d/output/spect_energy_budget.py:23-80
 *   2 ns3d spectra     (vx, vy, vz)               Q=5  vx vy vz Khr Khd
 *     fluidsim/solvers/ns3d/output/spectra.py:15-99
 *   3 strat spectra    (vx, vy, vz, b)            Q=6  vx vy vz A Khr Khd
 *     fluidsim/solvers/ns3d/strat/output/spectra.py:18-112
 *   4 ns3d budget      (v, P fft(v x omega))      Q=5  transfer_Kh transfer_Kz
 *                      transfer_Khr diss_Kh diss_Kz
 *     fluidsim/solvers/ns3d/output/spect_energy_budget.py:106-190
 *   5 strat budget     (v, f, b, b tendency)      Q=8  + diss_A conv_K2A transfer_A
 *     fluidsim/solvers/ns3d/strat/output/spect_energy_budget.py:60-118
 *   6 ns2d means       (rot [, Frot])             Q=10 E Z epsK epsK_hypo epsZ
 *                      epsZ_hypo PK1 PK2' PZ1 PZ2'
 *     fluidsim/solvers/ns2d/output/spatial_means.py:23-68
 *   7 ns3d means       (v [, f])                  Q=9  Ex Ey Ez epsK epsK_hypo
 *                      epsK4 epsK8 PK1 PK2'
 *     fluidsim/solvers/ns3d/output/spatial_means.py:23-77
 *   8 strat means      (v, b [, f, fb])           Q=18 A Kz Khr Khd Khs As epsK
 *                      epsK_hypo epsA epsA_hypo epsK4 epsA4 epsK8 epsA8 PK1 PK2'
 *                      PA1' PA2'
 *     fluidsim/solvers/ns3d/strat/output/spatial_means.py:27-106
 * (primed terms lack the factor dt/2 and, for PA, 1/N^2).  flags (spectra
 * and budgets only): 1 kx, 2 ky, 4 kz 1D spectra (bins |i_k|, / dk_i),
 * 8 shell spectrum (3D: dk = max dk_i, nk = max(2, int(min_i(dk_i (n_i/2+1))
 * / dk)); 2D: dk = dkx, nk = nx/2+1; floor bin and linear share with the
 * next bin, / dk), 16 kz-kh spectrum (kh bins as the 3D shell with dkx,
 * dky, / (dkz dkh)), replacing fluidfft's compute_1dspectra,
 * compute_3dspectrum, compute_spectrum_kzkh, spectra1D_from_fft and
 * spectrum2D_from_fft.  Each mode carries its Parseval weight.
 * fsx_diag_layout: out8 = Q, nkx, nky, nkz, nk (shell), nkh, total doubles,
 * required fields.  fsx_diag (sync) writes host_out[total] =
 * [kx Q*nkx][ky Q*nky][kz Q*nkz][shell Q*nk][kzkh Q*nkz*nkh][sums Q] (absent
 * parts take no room).  params = {1/N^2, nu_2, nu_4, nu_8, nu_m4}. */
int fsx_diag_layout(fsx_plan p, int kind, int flags, long long* out8);
int fsx_diag(fsx_plan p, int kind, int flags, const void* const* fields, int nfields,
             const double* params, double* host_out, void* stream);

/* ns2d budget nonlinear terms: out_fft (3 K fields) = dealiased fft of
 * Frot = -ux d_x rot - uy (d_y rot + beta), Fx = -ux d_x ux - uy d_y ux,
 * Fy = -ux d_x uy - uy d_y uy (u from rot), with caller workspace work_k
 * (8 K fields) and work_x (8 X fields)
 * (fluidsim/solvers/ns2d/output/spect_energy_budget.py:27-57). */
int fsx_ns2d_budget_fields(fsx_plan p, const void* rot_fft, double beta, void* work_k,
                           void* work_x, void* out_fft, void* stream);

/* ---- ensembles of independent ns2d runs ---------------------------------
 * No counterpart in the reference (one Simul per process there); the
 * per-run algorithm is the ns2d fused RK stage (fsx_ns2d_rk_stage,
 * fluidsim/solvers/ns2d/solver.py:111-194,
 * fluidsim/base/time_stepping/pseudo_spect.py:466-517, 798-984) at a fixed
 * dt per run.
 * fsx_plan_create_batch: a 2D plan for nbatch runs of one size; state
 * arrays hold the runs one after the other, (nbatch, ny, nx/2+1).
 * fsx_batch_set_params (sync): per-run dt[nbatch] and
 * visc4[4 * nbatch] = {nu_2, nu_4, nu_8, nu_m4} of each run.
 * fsx_batch_set_lines: lines per tile of the batched y-passes (power of two;
 * 0 = the launcher's choice), the knob the host autotuner sets.
 * fsx_ns2d_batch_rk_stage: one RK stage (scheme 4 or 2) of every run; no
 * CFL, no forcing, no allocation, capturable.
 * fsx_graph_begin / fsx_graph_end: capture the calls issued on `stream`
 * (a created stream, not the null stream) into an executable graph;
 * fsx_graph_launch replays it on a stream; fsx_graph_destroy frees it. */
int fsx_plan_create_batch(fsx_plan* out, int nx, int ny, double Lx, double Ly,
                          int precision, const unsigned char* where_dealiased,
                          int nbatch);
int fsx_batch_set_params(fsx_plan p, const double* dt, const double* visc4);
int fsx_batch_set_lines(fsx_plan p, int lines);
/* fsx_plan_set_tiles (no reference counterpart: tile choices of the 2D
 * y-passes of a single 2D plan, set by an autotuner): lines per tile of the
 * inverse and of the forward y-passes (powers of two; 0 = the launcher's
 * choice) and, with inv_split, the ns2d inverse y-pass with one output
 * field (u, v, d_x rot, d_y rot) per workgroup. */
int fsx_plan_set_tiles(fsx_plan p, int inv_lines, int fwd_lines, int inv_split);
/* fsx_batch_set_row: row pass of the batched stages, a choice of the
 * ensemble autotuner: 0 register rows with radix-8 stages (8 points per
 * lane in FP32; default), 1 register rows with radix-4 stages (FP32), 2 the
 * LDS row pass (Stockham radix 8/4/2 in LDS). */
int fsx_batch_set_row(fsx_plan p, int mode);
int fsx_ns2d_batch_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp, void* t1,
                            double beta, void* stream);
int fsx_graph_begin(void* stream);
int fsx_graph_end(void* stream, fsx_graph* out);
int fsx_graph_launch(fsx_graph g, void* stream);
int fsx_graph_destroy(fsx_graph g);

/* ---- convolutional subgrid closure (matrix cores) ------------------------
 * No counterpart in the reference (fluidsim's LES hook, TurbModel.get_forcing
 * in fluidsim/base/turb_model/base.py:87-88, has Smagorinsky only).
 * Channels-last layouts: activations BF16 [B][H][W][C] with C = 8 or 32
 * (zero-padded), pre-activations FP32 [B][H][W][co], weights BF16
 * [32][25][C] (tap t = 5 dy + dx, offset (dy - 2, dx - 2), periodic),
 * bias FP32 [32]; MFMA 32x32x16 (convolutions) and 32x32x8 (weight
 * gradients) BF16 with FP32 accumulation.
 * fsx_cnn_probe: D (32x32 FP32) = A (32x8 BF16) B (8x32 BF16), one wave.
 * fsx_cnn_probe16: the same with A (32x16) and B (16x32) through the
 *   K = 16 instruction the convolutions use.
 * fsx_cnn_conv_fwd: y = bias + periodic 5x5 correlation of x, co_out <= 32
 *   outputs; H multiple of 4, W multiple of 32, W <= 160 for C = 32
 *   (LDS of the 4-row tile; wider images return FSX_ERR_ARG).
 * fsx_cnn_conv_fwd_act: a hidden layer (32 outputs) with the activation in
 *   the epilogue: a = BF16(GELU(y)) (32 channels) and, when y is not NULL,
 *   the FP32 pre-activations y the backward pass needs.
 * fsx_cnn_conv_bwd_act: the input gradient of a hidden layer through its
 *   activation in one pass: d = BF16((conv(x, w) + bias) * GELU'(pre))
 *   (32 channels) with pre the layer's FP32 pre-activations, equal to
 *   fsx_cnn_conv_fwd into an FP32 buffer followed by fsx_cnn_act_bwd.
 * fsx_cnn_conv_flip: wt[c][24 - t][o] = w[o][t][c] (co_pad = 8 or 32
 *   input channels of wt), so fsx_cnn_conv_fwd(dy, wt) is the input gradient.
 * fsx_cnn_conv_wgrad: dw[o][t][c] = sum_p dy[p][o] x[p + d(t)][c] and
 *   dbias[o] = sum_p dy[p][o] (FP32), workspace from fsx_cnn_wgrad_workspace;
 *   deterministic (fixed-order reductions).
 * fsx_cnn_act: a = BF16(GELU(y)) (gelu = 1) or BF16(y), channel-padded.
 * fsx_cnn_act_bwd: d = BF16(g * GELU'(y)).
 * fsx_cnn_mse: loss = mean((out - target)^2) (host_loss: sync readback,
 *   NULL: none) and d = BF16(2 (out - target) / n) in channel 0 of cpad.
 * fsx_cnn_adam: Adam step of n FP32 values (bias-corrected, step >= 1),
 *   optional BF16 copy into wb.
 * fsx_cnn_adam_multi: fsx_cnn_adam on count (1 to 16) tensors in one
 *   launch; the arrays (host memory) hold each tensor's device pointers
 *   and size (wb NULL, or an array whose NULL entries skip the copy).
 * fsx_cnn_to_bf16: BF16 copy of n FP32 values.
 * fsx_cnn_forward: nl layers (GELU between them) in one call, inference;
 *   hidden layers write their BF16 activations from the convolution
 *   epilogue; work holds 2 x B H W x 32 BF16.
 * fsx_ns2d_closure_term: the learned subgrid term of a float32 ns2d plan in
 *   one call: c2r of rot_fft, the network (cin[0] = 8, cout[nl-1] = 1) on
 *   the vorticity, r2c into out_fft and dealiasing; the term enters the
 *   Runge-Kutta stages through fsx_ns2d_rk_stage_ext, as fluidsim adds a
 *   model term to the tendencies (fluidsim/solvers/ns3d/solver.py:246-249);
 *   workspace size from fsx_ns2d_closure_workspace. */
int fsx_cnn_probe(const void* A, const void* B, float* D, void* stream);
int fsx_cnn_probe16(const void* A, const void* B, float* D, void* stream);
int fsx_cnn_conv_fwd(const void* x, int B, int H, int W, int ci, const void* w,
                     const float* bias, int co_out, float* y, void* stream);
int fsx_cnn_conv_fwd_act(const void* x, int B, int H, int W, int ci, const void* w,
                         const float* bias, float* y, void* a, void* stream);
int fsx_cnn_conv_bwd_act(const void* x, int B, int H, int W, int ci, const void* w,
                         const float* bias, const float* pre, void* d, void* stream);
int fsx_cnn_conv_flip(const void* w, int ci, int co, int co_pad, void* wt, void* stream);
int fsx_cnn_wgrad_workspace(int B, int H, int ci, long long* bytes);
int fsx_cnn_conv_wgrad(const void* x, int ci, const void* dy, int co_in, int B, int H, int W,
                       float* dw, float* dbias, void* work, long long work_bytes,
                       void* stream);
/* fsx_cnn_conv_wgrad_noalloc: fsx_cnn_conv_wgrad (same kernels, same
 *   results bit for bit) with its partials in a static device array of the
 *   library instead of a workspace argument; it allocates nothing (no
 *   hipMalloc, no library allocation), for callers that cannot allocate:
 *   code under HIP graph capture, tight training loops, harnesses that
 *   forbid allocation.  No counterpart in the reference.  The array holds
 *   the largest partials buffer (256 x (32 x 25 x 32 + 32) floats, 26.2 MB
 *   per device, held until the process ends); sizes beyond it are refused.
 *   Every call on a device shares that device's array: calls that overlap on
 *   one device (other streams, other threads) race. */
int fsx_cnn_conv_wgrad_noalloc(const void* x, int ci, const void* dy, int co_in, int B, int H,
                               int W, float* dw, float* dbias, void* stream);
int fsx_cnn_act(const float* y, long long npix, int cin, int cpad, int use_gelu, void* a,
                void* stream);
int fsx_cnn_act_bwd(const float* g, int gs, const float* y, long long npix, int c, void* d,
                    void* stream);
int fsx_cnn_mse(const float* out, const float* target, long long npix, int cpad, void* d,
                double* work, double* host_loss, void* stream);
int fsx_cnn_adam(float* w, const float* g, float* m1, float* m2, int n, double lr, double b1,
                 double b2, double eps, int step, void* wb, void* stream);
int fsx_cnn_adam_multi(int count, float* const* w, const float* const* g, float* const* m1,
                       float* const* m2, const int* n, void* const* wb, double lr, double b1,
                       double b2, double eps, int step, void* stream);
int fsx_cnn_to_bf16(const float* in, int n, void* out, void* stream);
int fsx_cnn_forward(int nl, const int* cin, const int* cout, const void* const* w,
                    const float* const* b, const void* x, int B, int H, int W, float* y,
                    void* work, long long work_bytes, void* stream);
int fsx_ns2d_closure_workspace(fsx_plan p, long long* bytes);
int fsx_ns2d_closure_term(fsx_plan p, int nl, const int* cin, const int* cout,
                          const void* const* w, const float* const* b, const void* rot_fft,
                          void* out_fft, void* work, long long work_bytes, void* stream);

/* Live per-kernel timing (no reference counterpart: measurement support).
 * When enabled, HIP events are recorded around every pass launch on its
 * stream.  fsx_timing_read (sync) returns, per launch kind
 * (0 z-pass inverse, 1 y-pass inverse, 2 x row pass, 3 y-pass forward,
 * 4 z-pass forward, 5 Runge-Kutta stage, 6 other, 7 diagnostics pass of
 * fsx_diag), the summed
 * duration in ms, the summed algorithmic bytes (compulsory HBM reads +
 * writes of each launch) and the launch count; reset zeroes them. */
int fsx_timing_enable(fsx_plan p, int enable);
int fsx_timing_read(fsx_plan p, double* ms8, double* bytes8, long long* counts8,
                    int reset);
/* fsx_timing_read_network (sync): the same for the hidden 32 -> 32
 * convolutions of fsx_ns2d_closure_term issued with this plan, work in
 * flops (2 x pixels x 32 x 25 x 32 per launch); reset zeroes them (the
 * reset of fsx_timing_read does too). */
int fsx_timing_read_network(fsx_plan p, double* ms, double* flops, long long* count,
                            int reset);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDSIM_AMD_H */
