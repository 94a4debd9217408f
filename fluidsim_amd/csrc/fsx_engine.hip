// Plans, plain transforms, stage algebra, forcing, Smagorinsky, timing and
// graphs: the core C-ABI of the fluidsim_amd engine (gfx950).  The exported
// functions are declared in include/fluidsim_amd.h; every function returns 0
// on success or a negative code, and fsx_last_error() gives the message.
#include "fsx_internal.h"

thread_local std::string fsx_g_err;
thread_local fsx_plan_s* fsx_g_plan = nullptr;
thread_local int fsx_g_kind = K_OTHER;
thread_local double fsx_g_bytes = 0;

extern "C" {

int fsx_version(void) { return FSX_ABI_VERSION; }

int fsx_last_error(char* buf, int len) {
  if (!buf || len <= 0) return -1;
  std::snprintf(buf, (size_t)len, "%s", fsx_g_err.c_str());
  return 0;
}

int fsx_plan_create(fsx_plan* out, int dim, int nx, int ny, int nz, double Lx,
                    double Ly, double Lz, int precision,
                    const unsigned char* where_dealiased, int nf_max) {
  if (!out) return fail(FSX_ERR_ARG, "null plan pointer");
  *out = nullptr;
  if (dim != 2 && dim != 3) return fail(FSX_ERR_ARG, "dim must be 2 or 3");
  if (dim == 2) nz = 1;
  if (precision != 4 && precision != 8)
    return fail(FSX_ERR_ARG, "precision must be 4 (float32) or 8 (float64)");
  if (nx < 2 || ny < 2 || (dim == 3 && nz < 2))
    return fail(FSX_ERR_ARG, "grid dimensions must be >= 2");
  // power-of-two axes: register/LDS Stockham passes; other lengths:
  // mixed-radix passes with two LDS tile buffers (half the length)
  const int tile = (precision == 4 ? kEPT32 : kEPT64) * kThr;
  auto axis_max = [&](int n) { return is_pow2(n) ? tile : tile / 2; };
  if (ny > axis_max(ny) || nz > axis_max(nz) || (is_pow2(nx) ? nx / 2 > tile : nx > tile / 2))
    return fail(FSX_ERR_ARG, "grid dimension too large for the LDS tile");
  if (nf_max < 1 || nf_max > 8) return fail(FSX_ERR_ARG, "nf_max must be 1..8");
  if (!(Lx > 0) || !(Ly > 0) || (dim == 3 && !(Lz > 0)))
    return fail(FSX_ERR_ARG, "box lengths must be positive");

  fsx_plan_s* p = new fsx_plan_s();
  std::memset(p, 0, sizeof(*p));
  p->dim = dim;
  p->prec = precision;
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->nkx = nx / 2 + 1;
  p->Lx = Lx;
  p->Ly = Ly;
  p->Lz = dim == 3 ? Lz : 1.0;
  const double two_pi = 6.283185307179586476925286766559005768;
  p->dkx = two_pi / Lx;
  p->dky = two_pi / Ly;
  p->dkz = dim == 3 ? two_pi / Lz : 0.0;
  p->ymF = AxisMap{ny, ny / 2 + 1, ny - (ny / 2 + 1)};
  p->zmF = dim == 3 ? AxisMap{nz, nz / 2 + 1, nz - (nz / 2 + 1)} : AxisMap{1, 1, 0};
  p->mxF = p->nkx;
  p->nf_max = nf_max;
  // ns2d at ny >= 2048: one output field per workgroup in the y-inverse
  // (measured at 2048^2: 1979 -> 2177 RK4 steps/s; fsx_plan_set_tiles)
  p->ysplit = dim == 2 && ny >= 2048;

  // bounding box of the kept (non-dealiased) modes
  p->ymP = p->ymF;
  p->zmP = p->zmF;
  p->mxP = p->mxF;
  if (where_dealiased) {
    const int nkx = p->nkx;
    std::vector<char> kx_any(nkx, 0), ky_any(ny, 0), kz_any(nz, 0);
    size_t nkept = 0;
    for (int iz = 0; iz < nz; ++iz)
      for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nkx; ++ix)
          if (!where_dealiased[((size_t)iz * ny + iy) * nkx + ix]) {
            kx_any[ix] = ky_any[iy] = kz_any[iz] = 1;
            ++nkept;
          }
    int mx = 0;
    for (int ix = 0; ix < nkx; ++ix)
      if (kx_any[ix]) mx = ix + 1;
    auto axis = [](const std::vector<char>& any, int n) {
      int lo = 0, hi = 0;
      for (int i = 0; i <= n / 2; ++i)
        if (any[i]) lo = i + 1;
      for (int i = n - 1; i > n / 2; --i)
        if (any[i]) hi = n - i;
      return AxisMap{n, lo, hi};
    };
    if (nkept == 0) mx = 1;
    p->mxP = mx > 0 ? mx : 1;
    p->ymP = axis(ky_any, ny);
    if (p->ymP.nk() == 0) p->ymP = AxisMap{ny, 1, 0};
    p->zmP = dim == 3 ? axis(kz_any, nz) : AxisMap{1, 1, 0};
    if (p->zmP.nk() == 0) p->zmP = AxisMap{nz, 1, 0};
    const size_t box = (size_t)p->zmP.nk() * p->ymP.nk() * p->mxP;
    p->dealias = true;
    if (box != nkept) {
      const size_t n = (size_t)nz * ny * nkx;
      hipError_t e = hipMalloc((void**)&p->d_mask, n);
      if (e == hipSuccess)
        e = hipMemcpy(p->d_mask, where_dealiased, n, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        fsx_plan_destroy(p);
        return fail(FSX_ERR_HIP, std::string("mask upload: ") + hipGetErrorString(e));
      }
    }
  }

  int rc = 0;
  const int ns[3] = {nx, ny, nz};
  for (int a = 0; a < (dim == 3 ? 3 : 2) && !rc; ++a)
    rc = precision == 4 ? upload_twiddles<float>(ns[a], &p->tw[a])
                        : upload_twiddles<double>(ns[a], &p->tw[a]);
  if (rc) {
    fsx_plan_destroy(p);
    return rc;
  }
  const int align = 128 / (2 * precision);
  const size_t ldF = (size_t)((p->mxF + align - 1) / align * align);
  const size_t ldP = (size_t)((p->mxP + align - 1) / align * align);
  const size_t full = (size_t)nz * ny * ldF;
  const size_t w1p = (size_t)nf_max * nz * p->ymP.nk() * ldP;
  const size_t w2p = (size_t)nf_max * nz * ny * ldP;
  p->w1_elems = dim == 3 ? (w1p > full ? w1p : full) : 0;
  p->w2_elems = w2p > full ? w2p : full;
  const size_t csz = 2 * (size_t)precision;
  hipError_t e = hipSuccess;
  if (p->w1_elems) e = hipMalloc(&p->w1, p->w1_elems * csz);
  if (e == hipSuccess) e = hipMalloc(&p->w2, p->w2_elems * csz);
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_scal, 32 * sizeof(double));
  if (e == hipSuccess) e = hipMemset(p->d_scal, 0, 32 * sizeof(double));
  if (e != hipSuccess) {
    fsx_plan_destroy(p);
    return fail(FSX_ERR_HIP, std::string("workspace: ") + hipGetErrorString(e));
  }
  if ((rc = fsx_diag_alloc(p))) {
    fsx_plan_destroy(p);
    return rc;
  }
  *out = p;
  return 0;
}

int fsx_plan_destroy(fsx_plan p) {
  if (!p) return 0;
  if (p->tm) {
    for (hipEvent_t e : p->tm->ev) (void)hipEventDestroy(e);
    delete p->tm;
  }
  for (int a = 0; a < 3; ++a)
    if (p->tw[a]) (void)hipFree(p->tw[a]);
  if (p->w1) (void)hipFree(p->w1);
  if (p->w2) (void)hipFree(p->w2);
  if (p->d_mask) (void)hipFree(p->d_mask);
  if (p->d_scal) (void)hipFree(p->d_scal);
  if (p->d_fc) (void)hipFree(p->d_fc);
  if (p->d_fmodes) (void)hipFree(p->d_fmodes);
  if (p->d_fw) (void)hipFree(p->d_fw);
  if (p->d_ff0) (void)hipFree(p->d_ff0);
  if (p->d_ff1) (void)hipFree(p->d_ff1);
  if (p->d_dpart) (void)hipFree(p->d_dpart);
  if (p->d_dout) (void)hipFree(p->d_dout);
  if (p->d_bpar) (void)hipFree(p->d_bpar);
  delete p;
  return 0;
}

int fsx_plan_info(fsx_plan p, int* info, int n) {
  if (!p || !info) return fail(FSX_ERR_ARG, "null argument");
  const int v[12] = {p->dim, p->prec, p->nx, p->ny, p->nz, p->nkx,
                     p->mxP, p->ymP.lo, p->ymP.hi, p->zmP.lo, p->zmP.hi,
                     p->d_mask != nullptr};
  for (int i = 0; i < n && i < 12; ++i) info[i] = v[i];
  return 0;
}

int fsx_fft(fsx_plan p, const void* x, void* x_fft, int nfields, void* stream) {
  PlanScope scope_(p);
  if (!p || !x || !x_fft || nfields < 1) return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_fft<float>(p, x, x_fft, nfields, s)
                      : do_fft<double>(p, x, x_fft, nfields, s);
}

int fsx_ifft(fsx_plan p, const void* x_fft, void* x, int nfields, void* stream) {
  PlanScope scope_(p);
  if (!p || !x || !x_fft || nfields < 1) return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ifft<float>(p, x_fft, x, nfields, s)
                      : do_ifft<double>(p, x_fft, x, nfields, s);
}

int fsx_rk_stage(fsx_plan p, int scheme, int stage, int nvar, void* S,
                 void* tmp, void* t1, void* N, int project, int no_vz_kz0,
                 const double* visc, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !N || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  if (nvar < 1 || nvar > 4) return fail(FSX_ERR_ARG, "nvar must be 1..4");
  if (project && (p->dim != 3 || nvar < 3))
    return fail(FSX_ERR_ARG, "projection needs a 3D plan and 3 velocity fields");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_rk<float>(p, scheme, stage, nvar, S, tmp, t1, N,
                                     project, no_vz_kz0, visc, s)
                      : do_rk<double>(p, scheme, stage, nvar, S, tmp, t1, N,
                                      project, no_vz_kz0, visc, s);
}

int fsx_project(fsx_plan p, void* a, void* b, void* c, int kind, void* stream) {
  if (!p || !a || !b || !c || p->dim != 3) return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  const Box box = make_box(p, false);
  const size_t n = kelems(p);
  if (p->prec == 4)
    hipLaunchKernelGGL(k_project<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (cpx<float>*)a, (cpx<float>*)b, (cpx<float>*)c, kind, box);
  else
    hipLaunchKernelGGL(k_project<double>, dim3(grid_for(n)), dim3(256), 0, s,
                       (cpx<double>*)a, (cpx<double>*)b, (cpx<double>*)c, kind, box);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_dealias(fsx_plan p, void* arr, int nfields, int box_only, void* stream) {
  if (!p || !arr || nfields < 1) return fail(FSX_ERR_ARG, "bad argument");
  if (!p->dealias) return 0;
  hipStream_t s = (hipStream_t)stream;
  const Box box = make_box(p, true);
  const unsigned char* mask = box_only ? nullptr : p->d_mask;
  const size_t n = kelems(p);
  if (p->prec == 4)
    hipLaunchKernelGGL(k_dealias<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (cpx<float>*)arr, n, nfields, box, mask);
  else
    hipLaunchKernelGGL(k_dealias<double>, dim3(grid_for(n)), dim3(256), 0, s,
                       (cpx<double>*)arr, n, nfields, box, mask);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_set_dt(fsx_plan p, double dt, void* stream) {
  if (!p) return fail(FSX_ERR_ARG, "null plan");
  hipStream_t s = (hipStream_t)stream;
  static thread_local double h;
  h = dt;
  FSX_HIP(hipMemcpyAsync(p->d_scal + 4, &h, sizeof(double), hipMemcpyHostToDevice, s));
  FSX_HIP(hipStreamSynchronize(s));
  return 0;
}

int fsx_reset_maxima(fsx_plan p, void* stream) {
  if (!p) return fail(FSX_ERR_ARG, "null plan");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s,
                     reinterpret_cast<unsigned long long*>(p->d_scal), 4, 0ull);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cfl_update(fsx_plan p, int ncomp, double cfl, double dt_max,
                   void* stream) {
  if (!p || ncomp < 1 || ncomp > 3) return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  const double dx = p->Lx / p->nx, dy = p->Ly / p->ny,
               dz = p->dim == 3 ? p->Lz / p->nz : 1.0;
  hipLaunchKernelGGL(k_cfl_update, dim3(1), dim3(64), 0, s,
                     reinterpret_cast<const unsigned long long*>(p->d_scal),
                     ncomp, dx, dy, dz, cfl, dt_max, p->d_scal + 4);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_get_scalars(fsx_plan p, double* host6, void* stream) {
  if (!p || !host6) return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  double h[6];
  FSX_HIP(hipMemcpyAsync(h, p->d_scal, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
  FSX_HIP(hipStreamSynchronize(s));
  unsigned long long bits;
  for (int i = 0; i < 4; ++i) {
    std::memcpy(&bits, &h[i], 8);
    std::memcpy(&host6[i], &bits, 8);
  }
  host6[4] = h[4];
  unsigned int flag;
  std::memcpy(&flag, &h[5], 4);
  host6[5] = (double)flag;
  return 0;
}

int fsx_clear_nan_flag(fsx_plan p, void* stream) {
  if (!p) return fail(FSX_ERR_ARG, "null plan");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s,
                     reinterpret_cast<unsigned long long*>(p->d_scal + 5), 1, 0ull);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_sums(fsx_plan p, int kind, const void* f0, const void* f1,
             const void* f2, const void* f3, double N2, const double* visc,
             double* host_out6, void* stream) {
  if (!p || !f0 || !visc || !host_out6) return fail(FSX_ERR_ARG, "bad argument");
  if (kind != 0 && (!f1 || !f2)) return fail(FSX_ERR_ARG, "missing fields");
  if (kind == 2 && (!f3 || !(N2 > 0))) return fail(FSX_ERR_ARG, "strat needs b and N");
  hipStream_t s = (hipStream_t)stream;
  if (!f1) f1 = f0;
  if (!f2) f2 = f0;
  if (!f3) f3 = f0;
  return p->prec == 4
             ? do_sums<float>(p, kind, f0, f1, f2, f3, N2, visc, host_out6, s)
             : do_sums<double>(p, kind, f0, f1, f2, f3, N2, visc, host_out6, s);
}

int fsx_forcing_setup(fsx_plan p, int nyc, int nzc, int nkxc, int nvar,
                      int nmodes, const int* modes, const double* weights,
                      int key_kind, int key_var) {
  if (!p || nvar < 1 || nvar > 4 || nyc < 2 || nkxc < 2 || nmodes < 0 ||
      (nmodes > 0 && (!modes || !weights)))
    return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim == 2 && nzc != 1) return fail(FSX_ERR_ARG, "2D forcing box needs nzc = 1");
  if (p->dim == 3 && nzc < 2) return fail(FSX_ERR_ARG, "3D forcing box needs nzc >= 2");
  if (nyc > p->ny || nkxc > p->nkx || (p->dim == 3 && nzc > p->nz))
    return fail(FSX_ERR_ARG, "forcing box larger than the grid");
  if (key_kind != FKEY_VAR && key_kind != FKEY_VT) return fail(FSX_ERR_ARG, "bad key kind");
  if (key_kind == FKEY_VT && (p->dim != 3 || nvar < 3))
    return fail(FSX_ERR_ARG, "vt forcing needs a 3D plan and 3 velocity fields");
  if (key_kind == FKEY_VAR && (key_var < 0 || key_var >= nvar))
    return fail(FSX_ERR_ARG, "bad forced variable");
  const long long fvs = (long long)nzc * nyc * nkxc;
  for (int m = 0; m < nmodes; ++m)
    if (modes[m] < 0 || modes[m] >= fvs) return fail(FSX_ERR_ARG, "forced mode outside the box");
  const size_t csz = 2 * (size_t)p->prec;
  void* fc = nullptr;
  int* dm = nullptr;
  double* dw = nullptr;
  void *f0 = nullptr, *f1 = nullptr;
  hipError_t e = hipMalloc(&fc, fvs * nvar * csz);
  if (e == hipSuccess) e = hipMemset(fc, 0, fvs * nvar * csz);
  const size_t nm = nmodes > 0 ? nmodes : 1;
  if (e == hipSuccess) e = hipMalloc((void**)&dm, nm * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void**)&dw, nm * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&f0, nm * 2 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&f1, nm * 2 * sizeof(double));
  if (e == hipSuccess) e = hipMemset(f0, 0, nm * 2 * sizeof(double));
  if (e == hipSuccess) e = hipMemset(f1, 0, nm * 2 * sizeof(double));
  if (e == hipSuccess && nmodes > 0)
    e = hipMemcpy(dm, modes, nmodes * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && nmodes > 0)
    e = hipMemcpy(dw, weights, nmodes * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    for (void* q : {fc, (void*)dm, (void*)dw, f0, f1})
      if (q) (void)hipFree(q);
    return fail(FSX_ERR_HIP, std::string("forcing setup: ") + hipGetErrorString(e));
  }
  for (void* q : {p->d_fc, (void*)p->d_fmodes, (void*)p->d_fw, p->d_ff0, p->d_ff1})
    if (q) (void)hipFree(q);
  p->d_fc = fc;
  p->d_fmodes = dm;
  p->d_fw = dw;
  p->d_ff0 = f0;
  p->d_ff1 = f1;
  p->fvs = fvs;
  p->fnvar = nvar;
  p->fnyc = nyc;
  p->fnzc = nzc;
  p->fnkxc = nkxc;
  p->fnmodes = nmodes;
  p->fkey_kind = key_kind;
  p->fkey_var = key_var;
  p->force_pending = false;
  return 0;
}

int fsx_forcing_set_pair(fsx_plan p, const double* f0, const double* f1,
                         void* stream) {
  if (!p || !f0 || !f1) return fail(FSX_ERR_ARG, "bad argument");
  if (!p->d_ff0) return fail(FSX_ERR_ARG, "forcing box not set up (fsx_forcing_setup)");
  // earlier steps may still read the arrays on the stream
  FSX_HIP(hipStreamSynchronize((hipStream_t)stream));
  const size_t b = (size_t)p->fnmodes * 2 * sizeof(double);
  if (b) {
    FSX_HIP(hipMemcpy(p->d_ff0, f0, b, hipMemcpyHostToDevice));
    FSX_HIP(hipMemcpy(p->d_ff1, f1, b, hipMemcpyHostToDevice));
  }
  return 0;
}

int fsx_forcing_set_coarse(fsx_plan p, const void* fc, void* stream) {
  if (!p || !fc) return fail(FSX_ERR_ARG, "bad argument");
  if (!p->d_fc) return fail(FSX_ERR_ARG, "forcing box not set up (fsx_forcing_setup)");
  FSX_HIP(hipStreamSynchronize((hipStream_t)stream));
  FSX_HIP(hipMemcpy(p->d_fc, fc, p->fvs * p->fnvar * 2 * (size_t)p->prec,
                    hipMemcpyHostToDevice));
  return 0;
}

int fsx_forcing_enable(fsx_plan p, int on) {
  if (!p) return fail(FSX_ERR_ARG, "bad argument");
  if (on && !p->d_fc) return fail(FSX_ERR_ARG, "forcing box not set up (fsx_forcing_setup)");
  p->force_on = on != 0;
  return 0;
}

int fsx_forcing_step(fsx_plan p, double coef, double forcing_rate, int which_root) {
  if (!p) return fail(FSX_ERR_ARG, "bad argument");
  if (!p->d_fc || !p->d_fmodes || p->fnmodes <= 0)
    return fail(FSX_ERR_ARG, "forcing box not set up (fsx_forcing_setup)");
  if (which_root < 0 || which_root > 3) return fail(FSX_ERR_ARG, "bad root choice");
  p->fcoef = coef;
  p->frate = forcing_rate;
  p->froot = which_root;
  p->force_pending = true;
  return 0;
}

int fsx_forcing_graph(fsx_plan p, int on) {
  if (!p) return fail(FSX_ERR_ARG, "bad argument");
  p->fgraph = on != 0;
  return 0;
}

int fsx_forcing_push(fsx_plan p, void* stream) {
  if (!p) return fail(FSX_ERR_ARG, "bad argument");
  if (!p->fgraph) return fail(FSX_ERR_ARG, "the plan's forcing is not in graph mode");
  hipLaunchKernelGGL(k_forcing_par, dim3(1), dim3(64), 0, (hipStream_t)stream, p->d_scal + 16,
                     p->fcoef, p->frate, (double)p->froot);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_forcing_normalize(fsx_plan p, const void* S, void* stream) {
  PlanScope scope_(p);
  if (!p || !S) return fail(FSX_ERR_ARG, "bad argument");
  p->force_pending = false;
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? forcing_normalize<float>(p, S, s)
                      : forcing_normalize<double>(p, S, s);
}

int fsx_forcing_add(fsx_plan p, void* out_fft, int nvar, double scale, void* stream) {
  if (!p || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (!p->d_fc) return fail(FSX_ERR_ARG, "forcing box not set up (fsx_forcing_setup)");
  if (nvar < 1 || nvar > p->fnvar) return fail(FSX_ERR_ARG, "bad number of fields");
  hipStream_t s = (hipStream_t)stream;
  const long long ke = (long long)kelems(p);
  const int nz = p->dim == 3 ? p->nz : 1;
  const long long n = p->fvs * nvar;
  const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  if (p->prec == 4)
    hipLaunchKernelGGL(k_forcing_add<float>, dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<cpx<float>*>(out_fft), ke,
                       reinterpret_cast<const cpx<float>*>(p->d_fc), p->fvs, nvar,
                       p->fnyc, p->fnzc, p->fnkxc, p->ny, nz, p->nkx, (float)scale);
  else
    hipLaunchKernelGGL(k_forcing_add<double>, dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<cpx<double>*>(out_fft), ke,
                       reinterpret_cast<const cpx<double>*>(p->d_fc), p->fvs, nvar,
                       p->fnyc, p->fnzc, p->fnkxc, p->ny, nz, p->nkx, scale);
  FSX_HIP(hipGetLastError());
  return 0;
}


int fsx_strain_rate(fsx_plan p, const void* v_fft, void* S_fft, void* stream) {
  if (!p || !v_fft || !S_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3) return fail(FSX_ERR_ARG, "the strain rate needs a 3D plan");
  hipStream_t s = (hipStream_t)stream;
  const long long ke = (long long)kelems(p);
  if (p->prec == 4)
    hipLaunchKernelGGL(k_strain_rate<float>, dim3(grid_blocks(ke)), dim3(256), 0, s,
                       reinterpret_cast<const cpx<float>*>(v_fft),
                       reinterpret_cast<cpx<float>*>(S_fft), ke, kgrid(p));
  else
    hipLaunchKernelGGL(k_strain_rate<double>, dim3(grid_blocks(ke)), dim3(256), 0, s,
                       reinterpret_cast<const cpx<double>*>(v_fft),
                       reinterpret_cast<cpx<double>*>(S_fft), ke, kgrid(p));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_smagorinsky_phys(fsx_plan p, void* S, double coef, void* stream) {
  if (!p || !S) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3) return fail(FSX_ERR_ARG, "the Smagorinsky model needs a 3D plan");
  hipStream_t s = (hipStream_t)stream;
  const long long xe = (long long)xelems(p);
  if (p->prec == 4)
    hipLaunchKernelGGL(k_smag_phys<float>, dim3(grid_blocks(xe)), dim3(256), 0, s,
                       reinterpret_cast<float*>(S), xe, (float)coef);
  else
    hipLaunchKernelGGL(k_smag_phys<double>, dim3(grid_blocks(xe)), dim3(256), 0, s,
                       reinterpret_cast<double*>(S), xe, coef);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_smagorinsky_div(fsx_plan p, const void* g_fft, void* out_fft, int accumulate,
                        void* stream) {
  if (!p || !g_fft || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3) return fail(FSX_ERR_ARG, "the Smagorinsky model needs a 3D plan");
  hipStream_t s = (hipStream_t)stream;
  const long long ke = (long long)kelems(p);
  if (p->prec == 4)
    hipLaunchKernelGGL(k_smag_div<float>, dim3(grid_blocks(ke)), dim3(256), 0, s,
                       reinterpret_cast<const cpx<float>*>(g_fft),
                       reinterpret_cast<cpx<float>*>(out_fft), ke, kgrid(p), accumulate);
  else
    hipLaunchKernelGGL(k_smag_div<double>, dim3(grid_blocks(ke)), dim3(256), 0, s,
                       reinterpret_cast<const cpx<double>*>(g_fft),
                       reinterpret_cast<cpx<double>*>(out_fft), ke, kgrid(p), accumulate);
  FSX_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"

namespace {

template <typename T>
int if_combine(fsx_plan_s* p, int nvar, void* out, int nin, void* const* ins,
               const double* a, const double* b, const double* e, const int* pw,
               const double* shift, const double* visc, int check_nan, hipStream_t s) {
  IFCombine<T> c;
  for (int j = 0; j < 4; ++j) {
    const bool on = j < nin;
    c.in[j] = on ? reinterpret_cast<const cpx<T>*>(ins[j]) : nullptr;
    c.a[j] = on ? a[j] : 0.0;
    c.b[j] = on ? b[j] : 0.0;
    c.e[j] = on ? e[j] : 0.0;
    c.p[j] = on ? pw[j] : 0;
  }
  c.nin = nin;
  c.out = reinterpret_cast<cpx<T>*>(out);
  c.ke = (long long)kelems(p);
  c.nvar = nvar;
  c.ax = shift[0];
  c.ay = shift[1];
  c.az = shift[2];
  c.box = make_box(p, false);
  c.visc.nu2 = visc[0];
  c.visc.nu4 = visc[1];
  c.visc.nu8 = visc[2];
  c.visc.num4 = visc[3];
  c.dt = p->d_scal + 4;
  c.nanflag = check_nan ? reinterpret_cast<unsigned int*>(p->d_scal + 5) : nullptr;
  hipLaunchKernelGGL(k_if_combine<T>, dim3(grid_blocks(c.ke)), dim3(256), 0, s, c);
  FSX_HIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int fsx_if_combine(fsx_plan p, int nvar, void* out, int nin, void* const* ins,
                   const double* a, const double* b, const double* e,
                   const int* phase_pow, const double* shift, const double* visc,
                   int check_nan, void* stream) {
  if (!p || !out || !ins || !a || !b || !e || !phase_pow || !shift || !visc)
    return fail(FSX_ERR_ARG, "bad argument");
  if (nin < 1 || nin > 4 || nvar < 1 || nvar > 4) return fail(FSX_ERR_ARG, "bad sizes");
  for (int j = 0; j < nin; ++j) {
    if (!ins[j]) return fail(FSX_ERR_ARG, "null input");
    if (phase_pow[j] < -1 || phase_pow[j] > 1) return fail(FSX_ERR_ARG, "phase power must be -1, 0 or 1");
  }
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? if_combine<float>(p, nvar, out, nin, ins, a, b, e, phase_pow,
                                          shift, visc, check_nan, s)
                      : if_combine<double>(p, nvar, out, nin, ins, a, b, e, phase_pow,
                                           shift, visc, check_nan, s);
}

int fsx_resize_fft(const void* in, int nz_in, int ny_in, int nkx_in, void* out,
                   int nz_out, int ny_out, int nkx_out, int nfields, int precision,
                   void* stream) {
  if (!in || !out || nfields < 1 || nz_in < 1 || ny_in < 1 || nkx_in < 1 ||
      nz_out < 1 || ny_out < 1 || nkx_out < 1)
    return fail(FSX_ERR_ARG, "bad argument");
  if (precision != 4 && precision != 8) return fail(FSX_ERR_ARG, "precision must be 4 or 8");
  if ((nz_in == 1) != (nz_out == 1)) return fail(FSX_ERR_ARG, "2D and 3D arrays mixed");
  hipStream_t s = (hipStream_t)stream;
  const long long n = (long long)nz_out * ny_out * nkx_out * nfields;
  if (precision == 4)
    hipLaunchKernelGGL(k_resize<float>, dim3(grid_blocks(n)), dim3(256), 0, s,
                       reinterpret_cast<const cpx<float>*>(in), nz_in, ny_in, nkx_in,
                       reinterpret_cast<cpx<float>*>(out), nz_out, ny_out, nkx_out,
                       nfields);
  else
    hipLaunchKernelGGL(k_resize<double>, dim3(grid_blocks(n)), dim3(256), 0, s,
                       reinterpret_cast<const cpx<double>*>(in), nz_in, ny_in, nkx_in,
                       reinterpret_cast<cpx<double>*>(out), nz_out, ny_out, nkx_out,
                       nfields);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_timing_enable(fsx_plan p, int enable) {
  if (!p) return fail(FSX_ERR_ARG, "null plan");
  if (!p->tm) {
    p->tm = new fsx_timing();
    const int npairs = 2048;
    p->tm->ev.resize(2 * npairs);
    p->tm->kind.resize(npairs);
    p->tm->bytes.resize(npairs);
    for (auto& e : p->tm->ev) FSX_HIP(hipEventCreate(&e));
  }
  if (p->tm->on && !enable) {
    PlanScope scope_(p);
    int rc = tm_flush(p);
    if (rc) return rc;
  }
  p->tm->on = enable != 0;
  return 0;
}

int fsx_timing_read(fsx_plan p, double* ms, double* bytes, long long* counts,
                    int reset) {
  if (!p || !ms || !bytes || !counts) return fail(FSX_ERR_ARG, "bad argument");
  if (!p->tm) {
    for (int k = 0; k < fsx_timing::kReadKinds; ++k) ms[k] = bytes[k] = 0, counts[k] = 0;
    return 0;
  }
  int rc = tm_flush(p);
  if (rc) return rc;
  for (int k = 0; k < fsx_timing::kKinds; ++k) {
    if (k < fsx_timing::kReadKinds) {
      ms[k] = p->tm->ms[k];
      bytes[k] = p->tm->by[k];
      counts[k] = p->tm->cnt[k];
    }
    if (reset) {
      p->tm->ms[k] = p->tm->by[k] = 0;
      p->tm->cnt[k] = 0;
    }
  }
  return 0;
}

int fsx_timing_read_network(fsx_plan p, double* ms, double* flops, long long* count,
                            int reset) {
  if (!p || !ms || !flops || !count) return fail(FSX_ERR_ARG, "bad argument");
  *ms = *flops = 0;
  *count = 0;
  if (!p->tm) return 0;
  int rc = tm_flush(p);
  if (rc) return rc;
  *ms = p->tm->ms[K_CNN];
  *flops = p->tm->by[K_CNN];
  *count = p->tm->cnt[K_CNN];
  if (reset) {
    p->tm->ms[K_CNN] = p->tm->by[K_CNN] = 0;
    p->tm->cnt[K_CNN] = 0;
  }
  return 0;
}

}  // extern "C"

extern "C" {

int fsx_plan_create_batch(fsx_plan* out, int nx, int ny, double Lx, double Ly,
                          int precision, const unsigned char* where_dealiased,
                          int nbatch) {
  if (!out) return fail(FSX_ERR_ARG, "null plan pointer");
  *out = nullptr;
  if (nbatch < 1) return fail(FSX_ERR_ARG, "nbatch must be >= 1");
  fsx_plan p = nullptr;
  int rc = fsx_plan_create(&p, 2, nx, ny, 1, Lx, Ly, 1.0, precision, where_dealiased, 4);
  if (rc) return rc;
  // the batched passes use W2 for 4 fields of every run
  const size_t need = p->w2_elems * (size_t)nbatch;
  (void)hipFree(p->w2);
  p->w2 = nullptr;
  hipError_t e = hipMalloc(&p->w2, need * 2 * (size_t)precision);
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_bpar, 5 * sizeof(double) * nbatch);
  if (e == hipSuccess) e = hipMemset(p->d_bpar, 0, 5 * sizeof(double) * nbatch);
  if (e != hipSuccess) {
    fsx_plan_destroy(p);
    return fail(FSX_ERR_HIP, std::string("batch workspace: ") + hipGetErrorString(e));
  }
  p->w2_elems = need;
  p->nb = nbatch;
  *out = p;
  return 0;
}

int fsx_batch_set_params(fsx_plan p, const double* dt, const double* visc4) {
  if (!p || !dt || !visc4) return fail(FSX_ERR_ARG, "bad argument");
  if (p->nb < 1 || !p->d_bpar) return fail(FSX_ERR_ARG, "not a batch plan");
  std::vector<double> h(5 * (size_t)p->nb);
  for (int r = 0; r < p->nb; ++r) {
    if (!(dt[r] > 0)) return fail(FSX_ERR_ARG, "every run needs dt > 0");
    h[5 * r] = dt[r];
    for (int j = 0; j < 4; ++j) h[5 * r + 1 + j] = visc4[4 * r + j];
  }
  FSX_HIP(hipMemcpy(p->d_bpar, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int fsx_batch_set_lines(fsx_plan p, int lines) {
  if (!p || p->nb < 1) return fail(FSX_ERR_ARG, "not a batch plan");
  if (lines < 0 || (lines & (lines - 1))) return fail(FSX_ERR_ARG, "lines must be 0 or a power of two");
  p->blines = lines;
  return 0;
}


int fsx_plan_set_tiles(fsx_plan p, int inv_lines, int fwd_lines, int inv_split) {
  if (!p) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2 || p->nb >= 1) return fail(FSX_ERR_ARG, "a single 2D plan is needed");
  for (int l : {inv_lines, fwd_lines})
    if (l < 0 || (l & (l - 1))) return fail(FSX_ERR_ARG, "lines must be 0 or a power of two");
  p->ylines_inv = inv_lines;
  p->ylines_fwd = fwd_lines;
  p->ysplit = inv_split != 0;
  return 0;
}

int fsx_batch_set_row(fsx_plan p, int mode) {
  if (!p) return fail(FSX_ERR_ARG, "bad argument");
  if (p->nb < 1) return fail(FSX_ERR_ARG, "not a batch plan");
  if (mode < 0 || mode > 2) return fail(FSX_ERR_ARG, "row mode must be 0, 1 or 2");
  p->brow = mode;
  return 0;
}

struct fsx_graph_s {
  hipGraphExec_t exec;
};

int fsx_graph_begin(void* stream) {
  if (!stream) return fail(FSX_ERR_ARG, "capture needs a created stream (not the null stream)");
  FSX_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return 0;
}

int fsx_graph_end(void* stream, fsx_graph* out) {
  if (!stream || !out) return fail(FSX_ERR_ARG, "bad argument");
  *out = nullptr;
  hipGraph_t g = nullptr;
  FSX_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(FSX_ERR_HIP, std::string("graph: ") + hipGetErrorString(e));
  fsx_graph_s* h = new fsx_graph_s();
  h->exec = exec;
  *out = h;
  return 0;
}

int fsx_graph_launch(fsx_graph g, void* stream) {
  if (!g) return fail(FSX_ERR_ARG, "null graph");
  FSX_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return 0;
}

int fsx_graph_destroy(fsx_graph g) {
  if (!g) return 0;
  (void)hipGraphExecDestroy(g->exec);
  delete g;
  return 0;
}

}  // extern "C"
