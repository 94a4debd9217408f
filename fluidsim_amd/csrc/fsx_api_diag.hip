// C-ABI of the on-device diagnostics (spectra, spectral energy budgets,
// spatial means) and of the ns2d budget fields.
#include "fsx_internal.h"

namespace {

// LDS of one block record is kept within 128 KiB (one 512-thread block per
// CU): the five quantities of a 1024^3 spectrum fit one record, so the
// input is read once
constexpr int kDiagLdsDoubles = 16384;

// histogram copies per block (lane parity spreads same-bin adds) and the
// LDS budget of a block record with its copies (two blocks per CU; larger
// records get no copies)
constexpr int kDiagCopies = 4;
constexpr size_t kDiagLdsBytes = 80 * 1024;

int diag_geom(const fsx_plan_s* p, int kind, int flags, DiagGeom& g, DiagOut& o,
              long long& total) {
  if (kind < 0 || kind >= DIAG_NKINDS) return fail(FSX_ERR_ARG, "unknown diagnostics kind");
  const bool means = kind >= DIAG_NS2D_MEANS;
  const bool kind2d = kind == DIAG_NS2D_SPECTRA || kind == DIAG_NS2D_BUDGET ||
                      kind == DIAG_NS2D_MEANS;
  if (kind2d != (p->dim == 2))
    return fail(FSX_ERR_ARG, "diagnostics kind does not match the plan dimension");
  if (flags & ~31) return fail(FSX_ERR_ARG, "unknown histogram flags");
  if (means && flags) return fail(FSX_ERR_ARG, "the spatial-means kinds take no histograms");
  if (p->dim == 2 && (flags & (DIAG_KZ | DIAG_KZKH)))
    return fail(FSX_ERR_ARG, "kz histograms need a 3D plan");
  g = DiagGeom{};
  g.dim = p->dim;
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->nz;
  g.nkx = p->nkx;
  g.dkx = p->dkx;
  g.dky = p->dky;
  g.dkz = p->dim == 3 ? p->dkz : 1.0;
  g.nkxs = p->nx / 2 + 1;
  g.nkys = p->ny / 2 + 1;
  g.nkzs = p->dim == 3 ? p->nz / 2 + 1 : 1;
  if (p->dim == 3) {
    // k_spectra3d and kh_spectra of the 3D operators
    g.dk3 = std::max(p->dkx, std::max(p->dky, p->dkz));
    const double kmax = std::min(p->dkx * g.nkxs, std::min(p->dky * g.nkys, p->dkz * g.nkzs));
    g.nk3 = std::max(2, (int)(kmax / g.dk3));
    g.dkh = std::max(p->dkx, p->dky);
    const double khmax = std::min(p->dkx * g.nkxs, p->dky * g.nkys);
    g.nkh = std::max(2, (int)(khmax / g.dkh));
  } else {
    // khE = kxE of the 2D operators
    g.dk3 = p->dkx;
    g.nk3 = g.nkxs;
    g.dkh = p->dkx;
    g.nkh = 1;
  }
  g.flags = flags;
  const int Q = diag_nq(kind);
  o.Q = Q;
  long long off = 0;
  o.o_kx = off;
  if (flags & DIAG_KX) off += (long long)Q * g.nkxs;
  o.o_ky = off;
  if (flags & DIAG_KY) off += (long long)Q * g.nkys;
  o.o_kz = off;
  if (flags & DIAG_KZ) off += (long long)Q * g.nkzs;
  o.o_k3 = off;
  if (flags & DIAG_SHELL) off += (long long)Q * g.nk3;
  o.o_kzkh = off;
  if (flags & DIAG_KZKH) off += (long long)Q * g.nkzs * g.nkh;
  o.o_sum = off;
  total = off + Q;
  // blocks: one kz plane and a run of ky rows each, about 512 blocks
  int nchunk = (512 + p->nz - 1) / p->nz;
  nchunk = std::max(1, std::min(nchunk, std::max(1, p->ny / 4)));
  g.rows = (p->ny + nchunk - 1) / nchunk;
  g.nchunk = (p->ny + g.rows - 1) / g.rows;
  return 0;
}

int diag_record(DiagGeom& g, int q0, int nq) {
  g.q0 = q0;
  g.nq = nq;
  g.off_kx = 0;
  g.off_ky = g.off_kx + ((g.flags & DIAG_KX) ? nq * g.nkxs : 0);
  g.off_k3 = g.off_ky + ((g.flags & DIAG_KY) ? nq * g.nkys : 0);
  g.off_kh = g.off_k3 + ((g.flags & DIAG_SHELL) ? nq * g.nk3 : 0);
  g.off_sum = g.off_kh + ((g.flags & DIAG_KZKH) ? nq * g.nkh : 0);
  g.rec = g.off_sum + nq;
  // copies of the shell / kh histograms while the block's LDS stays within
  // kDiagLdsBytes (two blocks per CU)
  const int span = g.off_sum - g.off_k3;
  g.ncopy = 1;
  if (span > 0)
    for (int c = kDiagCopies; c > 1; c >>= 1)
      if ((size_t)(g.rec + (c - 1) * span) * sizeof(double) <= kDiagLdsBytes) {
        g.ncopy = c;
        break;
      }
  return g.rec;
}

size_t diag_lds(const DiagGeom& g) {
  return (size_t)(g.rec + (g.ncopy - 1) * (g.off_sum - g.off_k3)) * sizeof(double);
}

int diag_per_quantity(const DiagGeom& g) {
  return ((g.flags & DIAG_KX) ? g.nkxs : 0) + ((g.flags & DIAG_KY) ? g.nkys : 0) +
         ((g.flags & DIAG_SHELL) ? g.nk3 : 0) + ((g.flags & DIAG_KZKH) ? g.nkh : 0) + 1;
}

template <typename T, int KIND>
int diag_launch(const DiagIn& in, const Box& box, const DiagGeom& g, double* part,
                hipStream_t s) {
  const unsigned nb = (unsigned)(g.nz * g.nchunk);
  // compulsory bytes: every input field read once, the block records written
  int nin = 0;
  for (int f = 0; f < 8; ++f) nin += in.f[f] != nullptr;
  note(K_DIAG, (double)nin * g.nz * g.ny * g.nkx * sizeof(cpx<T>) +
                   (double)nb * g.rec * sizeof(double));
  const size_t lds = diag_lds(g);
  // records over 64 KiB leave room for one block per CU: 512 threads then,
  // 256 otherwise (two or more blocks per CU)
  auto go = [&](auto kern, unsigned nt, size_t& lds_set) {
    if (lds > lds_set) {
      int rc = set_lds(kern, lds);
      if (rc) return rc;
      lds_set = lds;
    }
    int rc = tm_pre(s);
    if (rc) return rc;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(nt), lds, s, in, box, g, part);
    FSX_HIP(hipGetLastError());
    return tm_post(s);
  };
  static size_t lds_small = 0, lds_big = 0;
  if (lds > 64 * 1024) return go(k_diag<T, KIND, 512>, 512, lds_big);
  return go(k_diag<T, KIND, 256>, 256, lds_small);
}

template <typename T>
int diag_dispatch(int kind, const DiagIn& in, const Box& box, const DiagGeom& g,
                  double* part, hipStream_t s) {
  switch (kind) {
    case DIAG_NS2D_SPECTRA: return diag_launch<T, DIAG_NS2D_SPECTRA>(in, box, g, part, s);
    case DIAG_NS2D_BUDGET: return diag_launch<T, DIAG_NS2D_BUDGET>(in, box, g, part, s);
    case DIAG_NS3D_SPECTRA: return diag_launch<T, DIAG_NS3D_SPECTRA>(in, box, g, part, s);
    case DIAG_STRAT_SPECTRA: return diag_launch<T, DIAG_STRAT_SPECTRA>(in, box, g, part, s);
    case DIAG_NS3D_BUDGET: return diag_launch<T, DIAG_NS3D_BUDGET>(in, box, g, part, s);
    case DIAG_STRAT_BUDGET: return diag_launch<T, DIAG_STRAT_BUDGET>(in, box, g, part, s);
    case DIAG_NS2D_MEANS: return diag_launch<T, DIAG_NS2D_MEANS>(in, box, g, part, s);
    case DIAG_NS3D_MEANS: return diag_launch<T, DIAG_NS3D_MEANS>(in, box, g, part, s);
    case DIAG_STRAT_MEANS: return diag_launch<T, DIAG_STRAT_MEANS>(in, box, g, part, s);
    default: return fail(FSX_ERR_ARG, "unknown diagnostics kind");
  }
}

// Block records and output of fsx_diag for one kind and flag set: the
// quantities of a launch (qstep) and the sizes in doubles.
int diag_sizes(const fsx_plan_s* p, int kind, int flags, size_t& part, size_t& out,
               int& qstep) {
  DiagGeom g;
  DiagOut o;
  long long total;
  int rc = diag_geom(p, kind, flags, g, o, total);
  if (rc) return rc;
  const int per_q = diag_per_quantity(g);
  if (per_q > kDiagLdsDoubles) return fail(FSX_ERR_ARG, "grid too large for the LDS histograms");
  qstep = std::max(1, std::min(o.Q, kDiagLdsDoubles / per_q));
  const size_t nb = (size_t)g.nz * g.nchunk;
  part = nb * (size_t)diag_record(g, 0, qstep);
  out = (size_t)total;
  return 0;
}

template <typename T>
int ns2d_budget(fsx_plan_s* p, const void* rot, double beta, void* wk, void* wx, void* out,
                hipStream_t s) {
  const long long ke = (long long)kelems(p), xe = (long long)xelems(p);
  hipLaunchKernelGGL(k_ns2d_budget_grads<T>, dim3(grid_blocks(ke)), dim3(256), 0, s,
                     reinterpret_cast<const cpx<T>*>(rot), reinterpret_cast<cpx<T>*>(wk), ke,
                     p->nkx, p->ny, p->dkx, p->dky);
  FSX_HIP(hipGetLastError());
  int rc = do_ifft<T>(p, wk, wx, 8, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ns2d_budget_prod<T>, dim3(grid_blocks(xe)), dim3(256), 0, s,
                     reinterpret_cast<T*>(wx), xe, beta);
  FSX_HIP(hipGetLastError());
  return do_fft<T>(p, wx, out, 3, s);
}

}  // namespace

// Buffers of fsx_diag, allocated with the plan: sized for the largest block
// records and output over the diagnostics kinds of the plan's dimension and
// every histogram set the grid admits (a large grid may admit only some of
// the histograms at once; sets it refuses are skipped).  A plan whose grid
// admits none gets no buffers, and fsx_diag refuses it.
int fsx_diag_alloc(fsx_plan_s* p) {
  size_t part = 0, out = 0;
  const std::string err = fsx_g_err;  // refused sets are not errors here
  for (int kind = 0; kind < DIAG_NKINDS; ++kind) {
    const bool kind2d = kind == DIAG_NS2D_SPECTRA || kind == DIAG_NS2D_BUDGET ||
                        kind == DIAG_NS2D_MEANS;
    if (kind2d != (p->dim == 2)) continue;
    const bool means = kind >= DIAG_NS2D_MEANS;
    const int all = means ? 0 : (p->dim == 3 ? 31 : 31 & ~(DIAG_KZ | DIAG_KZKH));
    for (int flags = 0; flags <= all; ++flags) {
      if ((flags & all) != flags) continue;
      size_t pa = 0, ou = 0;
      int qstep = 1;
      if (diag_sizes(p, kind, flags, pa, ou, qstep)) continue;
      part = std::max(part, pa);
      out = std::max(out, ou);
    }
  }
  fsx_g_err = err;
  if (!part || !out) return 0;
  FSX_HIP(hipMalloc((void**)&p->d_dpart, part * sizeof(double)));
  p->dpart_cap = part;
  FSX_HIP(hipMalloc((void**)&p->d_dout, out * sizeof(double)));
  p->dout_cap = out;
  return 0;
}

extern "C" {

int fsx_diag_layout(fsx_plan p, int kind, int flags, long long* out8) {
  if (!p || !out8) return fail(FSX_ERR_ARG, "bad argument");
  DiagGeom g;
  DiagOut o;
  long long total;
  int rc = diag_geom(p, kind, flags, g, o, total);
  if (rc) return rc;
  const int nin_req = kind == DIAG_NS2D_MEANS   ? 1
                      : kind == DIAG_NS3D_MEANS ? 3
                      : kind == DIAG_STRAT_MEANS ? 4
                                                 : diag_nin(kind);
  out8[0] = o.Q;
  out8[1] = g.nkxs;
  out8[2] = g.nkys;
  out8[3] = g.nkzs;
  out8[4] = g.nk3;
  out8[5] = g.nkh;
  out8[6] = total;
  out8[7] = nin_req;
  return 0;
}

int fsx_diag(fsx_plan p, int kind, int flags, const void* const* fields, int nfields,
             const double* params, double* host_out, void* stream) {
  PlanScope scope_(p);
  if (!p || !fields || !params || !host_out) return fail(FSX_ERR_ARG, "bad argument");
  DiagGeom g;
  DiagOut o;
  long long total;
  int rc = diag_geom(p, kind, flags, g, o, total);
  if (rc) return rc;
  const int nin = diag_nin(kind);
  const int nreq = kind == DIAG_NS2D_MEANS   ? 1
                   : kind == DIAG_NS3D_MEANS ? 3
                   : kind == DIAG_STRAT_MEANS ? 4
                                              : nin;
  if (nfields < nreq || nfields > nin) return fail(FSX_ERR_ARG, "wrong number of fields");
  DiagIn in{};
  for (int f = 0; f < nin; ++f) {
    in.f[f] = f < nfields ? fields[f] : nullptr;
    if (f < nreq && !in.f[f]) return fail(FSX_ERR_ARG, "null field");
  }
  if (nfields > nreq) {
    // optional forcing fields come all together
    for (int f = nreq; f < nin; ++f)
      if (!in.f[f]) return fail(FSX_ERR_ARG, "incomplete forcing fields");
  }
  in.inv_N2 = params[0];
  in.visc.nu2 = params[1];
  in.visc.nu4 = params[2];
  in.visc.nu8 = params[3];
  in.visc.num4 = params[4];
  const int Q = o.Q;
  size_t need_part = 0, need_out = 0;
  int qstep = 1;
  rc = diag_sizes(p, kind, flags, need_part, need_out, qstep);
  if (rc) return rc;
  // the plan holds buffers for the largest record of its dimension
  if (need_part > p->dpart_cap || need_out > p->dout_cap || !p->d_dpart || !p->d_dout)
    return fail(FSX_ERR_ARG, "diagnostics buffers of the plan too small");
  hipStream_t s = (hipStream_t)stream;
  const Box box = make_box(p, false);
  for (int q0 = 0; q0 < Q; q0 += qstep) {
    DiagGeom gq = g;
    diag_record(gq, q0, std::min(qstep, Q - q0));
    rc = p->prec == 4 ? diag_dispatch<float>(kind, in, box, gq, p->d_dpart, s)
                      : diag_dispatch<double>(kind, in, box, gq, p->d_dpart, s);
    if (rc) return rc;
    // one wave per output value
    const long long nred = std::min<long long>((total + 3) / 4, 65536);
    hipLaunchKernelGGL(k_diag_reduce, dim3((unsigned)nred), dim3(256), 0, s, p->d_dpart,
                       gq, o, p->d_dout);
    FSX_HIP(hipGetLastError());
  }
  FSX_HIP(hipMemcpyAsync(host_out, p->d_dout, (size_t)total * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  FSX_HIP(hipStreamSynchronize(s));
  return 0;
}

int fsx_ns2d_budget_fields(fsx_plan p, const void* rot_fft, double beta, void* work_k,
                           void* work_x, void* out_fft, void* stream) {
  PlanScope scope_(p);
  if (!p || !rot_fft || !work_k || !work_x || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2) return fail(FSX_ERR_ARG, "the ns2d budget needs a 2D plan");
  hipStream_t s = (hipStream_t)stream;
  int rc = p->prec == 4 ? ns2d_budget<float>(p, rot_fft, beta, work_k, work_x, out_fft, s)
                        : ns2d_budget<double>(p, rot_fft, beta, work_k, work_x, out_fft, s);
  if (rc) return rc;
  return fsx_dealias(p, out_fft, 3, 0, stream);
}


}  // extern "C"

