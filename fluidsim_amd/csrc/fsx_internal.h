// Shared internals of the fluidsim_amd engine (gfx950): plan structure,
// launch helpers and the pass drivers used by the C-ABI translation units
// (fsx_engine.hip, fsx_api_*.hip).  Device buffers are plain pointers
// (complex = interleaved re/im of the plan precision), streams are
// hipStream_t passed as void*.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fluidsim_amd.h"
#include "fsx_fft.h"
#include "fsx_kernels.h"
#include "fsx_pointwise.h"
#include "fsx_rows.h"
#include "fsx_forcing.h"
#include "fsx_turb.h"
#include "fsx_diag.h"
#include "fsx_cnn.h"

using namespace fsx;

// State shared by the translation units of the engine (defined in
// fsx_engine.hip).
struct fsx_plan_s;
int fsx_diag_alloc(fsx_plan_s* p);  // fsx_api_diag.hip
extern thread_local std::string fsx_g_err;      // message of the last failure
extern thread_local fsx_plan_s* fsx_g_plan;     // plan of the running API call
extern thread_local int fsx_g_kind;             // timing kind of the next launch
extern thread_local double fsx_g_bytes;         // its algorithmic bytes

namespace {

int fail(int code, const std::string& msg) {
  fsx_g_err = msg;
  return code;
}

#define FSX_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return fail(FSX_ERR_HIP, std::string(#call) + ": " +                 \
                                   hipGetErrorString(e_));                 \
  } while (0)

int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}
bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

}  // namespace

// Live per-kernel timing: HIP events recorded around each pass launch on the
// stream it runs on, with the algorithmic bytes of that launch (compulsory
// HBM reads + writes of its inputs and outputs).
struct fsx_timing {
  static constexpr int kKinds = 9;     // kind 8: hidden-layer convolutions
  static constexpr int kReadKinds = 8;  // kinds returned by fsx_timing_read
  bool on = false;
  std::vector<hipEvent_t> ev;  // pairs
  std::vector<int> kind;
  std::vector<double> bytes;
  int used = 0;
  double ms[kKinds] = {0};
  double by[kKinds] = {0};
  long long cnt[kKinds] = {0};
};

struct fsx_plan_s {
  fsx_timing* tm;
  int dim, prec;
  int nx, ny, nz, nkx;
  double Lx, Ly, Lz, dkx, dky, dkz;
  AxisMap ymF, zmF;  // all modes
  int mxF;
  AxisMap ymP, zmP;  // bounding box of the non-dealiased modes
  int mxP;
  unsigned char* d_mask;  // where_dealiased when the kept set is not a box
  bool dealias;
  void* tw[3];  // x, y, z twiddle tables
  void* w1;
  void* w2;
  size_t w1_elems, w2_elems;  // capacity in complex elements
  int nf_max;
  double* d_scal;  // [0..3] maxima bits, [4] dt, [5] nan flag, [8..15] sums,
                   // [16..18] forcing coef, rate, root (graph mode)
  // coarse forcing box (fsx_forcing_*)
  bool force_on, force_pending;
  void* d_fc;
  long long fvs;
  int fnvar, fnyc, fnzc, fnkxc;
  int* d_fmodes;
  double* d_fw;
  void* d_ff0;
  void* d_ff1;
  int fnmodes;
  int fkey_kind, fkey_var;
  double fcoef, frate;
  int froot;
  // the normalisation reads coef / rate / root from d_scal[16..18]
  // (fsx_forcing_graph), pushed before each step by fsx_forcing_push
  bool fgraph;
  // diagnostics: per-block records and the output (fsx_diag)
  double* d_dpart;
  size_t dpart_cap;
  double* d_dout;
  size_t dout_cap;
  // batch of nb independent 2D runs (fsx_plan_create_batch): per-run
  // {dt, nu_2, nu_4, nu_8, nu_m4} and the lines-per-tile cap of the batched
  // y-passes (0: the launcher's choice)
  int nb;
  double* d_bpar;
  int blines;
  int brow;  // row pass of the batch (fsx_batch_set_row)
  // 2D y-pass tiles of a single plan (fsx_plan_set_tiles): lines per tile of
  // the inverse and forward y-passes (0: the launcher's choice) and the ns2d
  // inverse with one output field per workgroup
  int ylines_inv, ylines_fwd, ysplit;
};

namespace {

// kinds of launches reported by fsx_timing_read
enum { K_ZINV = 0, K_YINV = 1, K_XPASS = 2, K_YFWD = 3, K_ZFWD = 4, K_RK = 5,
       K_OTHER = 6, K_DIAG = 7, K_CNN = 8 };


struct PlanScope {
  explicit PlanScope(fsx_plan_s* p) { fsx_g_plan = p; }
  ~PlanScope() { fsx_g_plan = nullptr; }
};

void note(int kind, double bytes) {
  fsx_g_kind = kind;
  fsx_g_bytes = bytes;
}

int tm_flush(fsx_plan_s* p);

int tm_pre(hipStream_t s) {
  fsx_plan_s* p = fsx_g_plan;
  if (!p || !p->tm || !p->tm->on) return 0;
  fsx_timing* t = p->tm;
  if ((size_t)(2 * t->used + 2) > t->ev.size()) {
    int rc = tm_flush(p);
    if (rc) return rc;
  }
  FSX_HIP(hipEventRecord(t->ev[2 * t->used], s));
  return 0;
}

int tm_post(hipStream_t s) {
  fsx_plan_s* p = fsx_g_plan;
  if (!p || !p->tm || !p->tm->on) return 0;
  fsx_timing* t = p->tm;
  FSX_HIP(hipEventRecord(t->ev[2 * t->used + 1], s));
  t->kind[t->used] = fsx_g_kind;
  t->bytes[t->used] = fsx_g_bytes;
  ++t->used;
  return 0;
}

int tm_flush(fsx_plan_s* p) {
  fsx_timing* t = p->tm;
  for (int i = 0; i < t->used; ++i) {
    FSX_HIP(hipEventSynchronize(t->ev[2 * i + 1]));
    float ms = 0;
    FSX_HIP(hipEventElapsedTime(&ms, t->ev[2 * i], t->ev[2 * i + 1]));
    const int k = t->kind[i];
    t->ms[k] += ms;
    t->by[k] += t->bytes[i];
    t->cnt[k] += 1;
  }
  t->used = 0;
  return 0;
}

constexpr int kThr = 512;   // threads per workgroup of the strided passes
constexpr int kXThr = 512;       // threads per workgroup of the LDS row pass
constexpr int kXPassEptDiv = 2;  // its tile: half the strided passes' elements per thread
constexpr int kEPT32 = 16;  // 8192 complex64 per tile  = 64 KiB of LDS
constexpr int kEPT64 = 8;   // 4096 complex128 per tile = 64 KiB of LDS
constexpr int kMinBlocks = 1024;  // shrink tiles of small problems below this

template <typename T>
constexpr int ept() {
  return sizeof(T) == 4 ? kEPT32 : kEPT64;
}

// row stride rounded up to 128 bytes
template <typename T>
int padded(int m) {
  const int a = 128 / (int)sizeof(cpx<T>);
  return (m + a - 1) / a * a;
}

template <typename T>
int upload_twiddles(int n, void** out) {
  std::vector<cpx<T>> h(n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    const long double a = two_pi * (long double)m / (long double)n;
    h[m].x = (T)cosl(a);
    h[m].y = (T)(-sinl(a));
  }
  FSX_HIP(hipMalloc(out, n * sizeof(cpx<T>)));
  FSX_HIP(hipMemcpy(*out, h.data(), n * sizeof(cpx<T>), hipMemcpyHostToDevice));
  return 0;
}

constexpr size_t kLdsMax = 160 * 1024;  // LDS per workgroup on gfx950

template <typename K>
int set_lds(K kern, size_t lds) {
  FSX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  return 0;
}

// Lines per tile an op may transform at once, as a multiple of L (an op
// with several outputs per line can stack them as extra lines).
template <class Op, class = void>
struct line_mul { static constexpr int value = 1; };
template <class Op>
struct line_mul<Op, std::void_t<decltype(Op::kLineMul)>> {
  static constexpr int value = Op::kLineMul;
};

// LDS of a strided tile of `tile` elements (N positions x its lines; the
// kernel touches only the padded tile after the N twiddles)
template <typename T, int DIR, int E, int KT, class Op, bool GEN = false>
int launch_strided_tile(Op& op, int N, size_t tile_max, long long blocks, int nfb,
                        hipStream_t s) {
  const size_t tile = std::min(tile_max, (size_t)E * KT);
  // mixed-radix lengths: two tile buffers (gen_fft stages are out of place)
  const size_t lds = ((GEN ? 2 : 1) * (tile + tile / 32 + 1) + N) * sizeof(cpx<T>);
  auto kern = k_strided<T, DIR, E, KT, Op, GEN>;
  static size_t lds_set = 0;
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "pass needs more LDS than a workgroup has");
  if (lds > lds_set) {
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    lds_set = lds;
  }
  int rc = tm_pre(s);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks, (unsigned)nfb), dim3(KT), lds, s, op);
  FSX_HIP(hipGetLastError());
  return tm_post(s);
}

template <typename T, int DIR, class Op>
int launch_strided(Op& op, int N, int nouter, int nlines, int nf, const void* tw,
                   hipStream_t s, bool kio = false, int lines_cap = 0) {
  constexpr int E = ept<T>() / Op::kEptDiv >= 8 ? ept<T>() / Op::kEptDiv : 8;
  const int tile = E * kThr;
  if (N > tile) return fail(FSX_ERR_ARG, "transform length exceeds the tile");
  const int nfb = Op::kFieldPerBlock ? nf : 1;
  // FP32 transforms whose tile of kThr threads holds short runs per
  // position run 2 kThr threads per workgroup (twice the lines) when that
  // tile's LDS fits: the multi-transform z-passes (halved tiles) when runs
  // are shorter than a 128-byte line, plain passes below 8 lines (measured:
  // wider plain y-passes at 8 lines are slower)
  constexpr int KW = 2 * kThr;
  constexpr int wide_below = Op::kEptDiv > 1 ? (int)(128 / sizeof(cpx<T>)) : 8;
  // lengths that are not powers of two: mixed-radix stages in LDS, 512-thread
  // tiles of at most half the size (two buffers)
  const bool gen = !is_pow2(N);
  const bool wide_ok = !gen && sizeof(T) == 4 && tile / N < wide_below &&
                       ((size_t)E * KW + E * KW / 32 + N) * sizeof(cpx<T>) <= kLdsMax;
  int L = (wide_ok ? E * KW : tile) / N;
  if (gen) {
    op.d.nst = gen_factor(N, op.d.rad);
    if (op.d.nst < 0) return fail(FSX_ERR_ARG, "transform length has too many factors");
    L = (tile / 2) / N;
    if (L < 1) return fail(FSX_ERR_ARG, "transform length exceeds the tile");
  } else {
    op.d.nst = 0;
  }
  while (L & (L - 1)) L &= L - 1;  // power of two
  int cap = 1;
  while (cap < nlines) cap <<= 1;
  if (L > cap) L = cap;
  if (lines_cap > 0) {
    // tuned lines per tile (power of two)
    while (L > lines_cap && L > 1) L >>= 1;
  } else {
    // enough workgroups to fill the chip when the problem is small; long
    // transforms keep more lines per tile (measured on the 2D y-passes:
    // at length 2048, 4 lines in 171 workgroups beat 1 line in 683; at
    // length 1024, 1 line in 342 beats 2 in 171)
    const int min_blocks = N >= 2048 ? 128 : kMinBlocks;
    while (L > 1 && (long long)nfb * nouter * ((nlines + L - 1) / L) < min_blocks) L >>= 1;
  }
  const bool wide = wide_ok && L > tile / N;  // a kThr tile cannot hold L lines
  op.d.N = N;
  op.d.logN = ilog2(N);
  op.d.L = L;
  op.d.logL = ilog2(L);
  op.d.nouter = nouter;
  op.d.nlines = nlines;
  op.d.nf = nf;
  op.d.tw = reinterpret_cast<const cpx<T>*>(tw);
  // tiles whose runs are shorter than a 128-byte cache line share lines with
  // their neighbours: keep neighbours on one XCD as for K rows
  op.d.kio = kio || (size_t)L * sizeof(cpx<T>) < 128;
  if (nouter <= 0 || nlines <= 0 || nf <= 0) return 0;
  const long long nchunks = (nlines + L - 1) / L;
  const long long blocks = nchunks * nouter;
  if (blocks > 0x7fffffffLL) return fail(FSX_ERR_ARG, "grid too large");
  // elements the op transforms at once: N L, or N L times its line
  // multiple when that fits a kThr tile
  size_t need = (size_t)N * L;
  if (gen) {
    // the same test as the op's own (stacked lines fit the kThr tile)
    if (need * line_mul<Op>::value <= (size_t)E * kThr) need *= line_mul<Op>::value;
    return launch_strided_tile<T, DIR, E, kThr, Op, true>(op, N, need, blocks, nfb, s);
  }
  if (!wide && need * line_mul<Op>::value <= (size_t)E * kThr) need *= line_mul<Op>::value;
  // 1024-thread tiles are FP32 only (wide_ok): no FP64 instantiation
  if constexpr (sizeof(T) == 4)
    if (wide) return launch_strided_tile<T, DIR, E, KW>(op, N, need, blocks, nfb, s);
  // small problems (few lines per tile): as many threads as the tile needs
  if ((size_t)E * kThr / 4 >= need)
    return launch_strided_tile<T, DIR, E, kThr / 4>(op, N, need, blocks, nfb, s);
  if ((size_t)E * kThr / 2 >= need)
    return launch_strided_tile<T, DIR, E, kThr / 2>(op, N, need, blocks, nfb, s);
  return launch_strided_tile<T, DIR, E, kThr>(op, N, need, blocks, nfb, s);
}

// Row pass of a length nx that is not a power of two (k_xpass_gen).
template <typename T, class Op>
int launch_xpass_gen(Op& op, int nx, int nrows, int nf_in, int nf_out, int mx_in,
                     int ld_in, int mx_out, int ld_out, long long ifs, long long ofs,
                     T scl, const void* tw, hipStream_t s) {
  constexpr int E = ept<T>() / kXPassEptDiv >= 8 ? ept<T>() / kXPassEptDiv : 8;
  (void)E;
  // two buffers of complex rows in the workgroup's LDS, after the twiddles
  const int tile = (int)((kLdsMax / sizeof(cpx<T>) - (size_t)nx) / 2);
  const int nfm = nf_in > nf_out ? nf_in : nf_out;
  const int ld = nx + 1;
  if (nfm * ld > tile) return fail(FSX_ERR_ARG, "row pass exceeds the tile");
  int RB = tile / (nfm * ld);
  if (RB > nrows) RB = nrows > 0 ? nrows : 1;
  while (RB > 1 && (nrows + RB - 1) / RB < kMinBlocks) RB = (RB + 1) / 2;
  op.d.nx = nx;
  op.d.lognx = -1;
  op.d.M = 0;
  op.d.logM = -1;
  op.d.RB = RB;
  op.d.logRB = -1;
  op.d.nrows = nrows;
  op.d.nf_in = nf_in;
  op.d.nf_out = nf_out;
  op.d.mx_in = mx_in;
  op.d.ld_in = ld_in;
  op.d.mx_out = mx_out;
  op.d.ld_out = ld_out;
  op.d.ifs = ifs;
  op.d.ofs = ofs;
  op.d.scale = scl;
  op.d.tw = reinterpret_cast<const cpx<T>*>(tw);
  op.d.nst = gen_factor(nx, op.d.rad);
  if (op.d.nst < 0) return fail(FSX_ERR_ARG, "row length has too many factors");
  if (nrows <= 0) return 0;
  const long long blocks = (nrows + RB - 1) / RB;
  const size_t lds = ((size_t)2 * nfm * RB * ld + nx) * sizeof(cpx<T>);
  auto kern = k_xpass_gen<T, kXThr, Op>;
  static size_t lds_set = 0;
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "pass needs more LDS than a workgroup has");
  if (lds > lds_set) {
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    lds_set = lds;
  }
  int rc = tm_pre(s);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kXThr), lds, s, op);
  FSX_HIP(hipGetLastError());
  return tm_post(s);
}

template <typename T, class Op>
int launch_xpass(Op& op, int nx, int nrows, int nf_in, int nf_out, int mx_in,
                 int ld_in, int mx_out, int ld_out, long long ifs,
                 long long ofs, T scl, const void* tw, hipStream_t s) {
  if (!is_pow2(nx) || nx < 4)
    return launch_xpass_gen<T>(op, nx, nrows, nf_in, nf_out, mx_in, ld_in, mx_out, ld_out,
                               ifs, ofs, scl, tw, s);
  constexpr int E = ept<T>() / kXPassEptDiv >= 8 ? ept<T>() / kXPassEptDiv : 8;
  const int tile = E * kXThr;
  const int M = nx / 2;
  const int nfm = nf_in > nf_out ? nf_in : nf_out;
  if (nfm * M > tile) return fail(FSX_ERR_ARG, "row pass exceeds the tile");
  int RB = 1;
  while (2 * RB * nfm * M <= tile && 2 * RB <= nrows) RB *= 2;
  while (RB > 1 && (nrows + RB - 1) / RB < kMinBlocks) RB >>= 1;
  op.d.nx = nx;
  op.d.lognx = ilog2(nx);
  op.d.M = M;
  op.d.logM = ilog2(M);
  op.d.RB = RB;
  op.d.logRB = ilog2(RB);
  op.d.nrows = nrows;
  op.d.nf_in = nf_in;
  op.d.nf_out = nf_out;
  op.d.mx_in = mx_in;
  op.d.ld_in = ld_in;
  op.d.mx_out = mx_out;
  op.d.ld_out = ld_out;
  op.d.ifs = ifs;
  op.d.ofs = ofs;
  op.d.scale = scl;
  op.d.tw = reinterpret_cast<const cpx<T>*>(tw);
  if (nrows <= 0) return 0;
  const long long blocks = (nrows + RB - 1) / RB;
  const size_t lds = ((size_t)tile + tile / 32 + nfm * RB + nx) * sizeof(cpx<T>);
  auto kern = k_xpass<T, E, kXThr, Op>;
  static size_t lds_set = 0;
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "pass needs more LDS than a workgroup has");
  if (lds > lds_set) {
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    lds_set = lds;
  }
  int rc = tm_pre(s);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kXThr), lds, s, op);
  FSX_HIP(hipGetLastError());
  return tm_post(s);
}

// Register-resident fused row pass (k_xrow): 2^LOGP points per lane
// (8 in FP32, 4 in FP64), a row within one wave.  Returns 1 when the row
// length is outside what it covers (the caller then uses k_xpass).
constexpr int kXRThr = 256;          // threads per workgroup
constexpr int kXRBlocksPerCU = 4;    // grid cap: workgroups per CU

template <typename T>
constexpr int xrow_logp() {
  return sizeof(T) == 4 ? 3 : 2;
}

template <typename T, int LOGM, class Op, int LOGP = xrow_logp<T>()>
int launch_xrow_m(const XRDesc<T>& d, const Op& op, hipStream_t s) {
  using PL = RowPlan<LOGM, LOGP>;
  constexpr int RPB = (kXRThr / 64) * (64 / PL::G);
  const size_t lds = ((size_t)RPB * xrow_ld<T>(PL::M) + 0) *
                     sizeof(cpx<T>);
  auto kern = k_xrow<T, LOGM, LOGP, kXRThr, Op>;
  static size_t lds_set = 0;
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "row pass needs more LDS than a workgroup has");
  if (lds > lds_set) {
    int rc = set_lds(kern, lds);
    if (rc) return rc;
    lds_set = lds;
  }
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    FSX_HIP(hipGetDevice(&dev));
    FSX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu <= 0) ncu = 256;
  }
  const long long need = (d.nrows + RPB - 1) / RPB;
  const long long cap = (long long)ncu * kXRBlocksPerCU;
  const unsigned blocks = (unsigned)(need < cap ? need : cap);
  if (d.nrows <= 0) return 0;
  int rc = tm_pre(s);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kXRThr), lds, s, d, op);
  FSX_HIP(hipGetLastError());
  return tm_post(s);
}

template <typename T, class Op>
int launch_xrow(int nx, int nrows, int mx_in, int ld_in, int mx_out, int ld_out,
                long long ifs, long long ofs, T scl, const void* in, void* out,
                const void* tw, unsigned long long* maxima, const Op& op,
                hipStream_t s, bool radix4 = false) {
  if (!is_pow2(nx)) return 1;
  XRDesc<T> d;
  d.nrows = nrows;
  d.mx_in = mx_in;
  d.ld_in = ld_in;
  d.mx_out = mx_out;
  d.ld_out = ld_out;
  d.ifs = ifs;
  d.ofs = ofs;
  d.scale = scl;
  d.in = reinterpret_cast<const cpx<T>*>(in);
  d.out = reinterpret_cast<cpx<T>*>(out);
  d.tw = reinterpret_cast<const cpx<T>*>(tw);
  d.maxima = maxima;
  if constexpr (sizeof(T) == 4 && std::is_same<Op, XRNS2D<T>>::value) {
    // ns2d rows with 4 points per lane (radix-4 stages), a choice of the
    // ensemble autotuner
    if (radix4) {
      switch (ilog2(nx / 2) - 2) {
        case 0: return launch_xrow_m<T, 2, Op, 2>(d, op, s);
        case 1: return launch_xrow_m<T, 3, Op, 2>(d, op, s);
        case 2: return launch_xrow_m<T, 4, Op, 2>(d, op, s);
        case 3: return launch_xrow_m<T, 5, Op, 2>(d, op, s);
        case 4: return launch_xrow_m<T, 6, Op, 2>(d, op, s);
        case 5: return launch_xrow_m<T, 7, Op, 2>(d, op, s);
        case 6: return launch_xrow_m<T, 8, Op, 2>(d, op, s);
        default: return 1;
      }
    }
  }
  constexpr int LOGP = xrow_logp<T>();
  switch (ilog2(nx / 2) - LOGP) {
    case 0: return launch_xrow_m<T, LOGP + 0>(d, op, s);
    case 1: return launch_xrow_m<T, LOGP + 1>(d, op, s);
    case 2: return launch_xrow_m<T, LOGP + 2>(d, op, s);
    case 3: return launch_xrow_m<T, LOGP + 3>(d, op, s);
    case 4: return launch_xrow_m<T, LOGP + 4>(d, op, s);
    case 5: return launch_xrow_m<T, LOGP + 5>(d, op, s);
    case 6: return launch_xrow_m<T, LOGP + 6>(d, op, s);
    default: return 1;  // longer rows: the LDS row pass (k_xpass)
  }
}

Box make_box(const fsx_plan_s* p, bool pruned) {
  Box b;
  b.dim = p->dim;
  b.nx = p->nx;
  b.ny = p->ny;
  b.nz = p->nz;
  b.nkx = p->nkx;
  b.mx = pruned ? p->mxP : p->mxF;
  b.ym = pruned ? p->ymP : p->ymF;
  b.zm = pruned ? p->zmP : p->zmF;
  b.dkx = p->dkx;
  b.dky = p->dky;
  b.dkz = p->dkz;
  return b;
}

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Workspace geometry of a pass set (pruned = solver passes, else plain
// transforms over all modes).
struct Geo {
  AxisMap ym, zm;
  int mx, ld1, ld2;
  long long fs1, fs2;  // field strides of W1 and W2 (elements)
};

template <typename T>
Geo geo(const fsx_plan_s* p, bool pruned) {
  Geo g;
  g.ym = pruned ? p->ymP : p->ymF;
  g.zm = pruned ? p->zmP : p->zmF;
  g.mx = pruned ? p->mxP : p->mxF;
  g.ld1 = padded<T>(g.mx);
  g.ld2 = padded<T>(g.mx);
  g.fs1 = (long long)p->nz * g.ym.nk() * g.ld1;
  g.fs2 = (long long)p->nz * p->ny * g.ld2;
  return g;
}

template <typename T>
cpx<T>* W1(const fsx_plan_s* p) {
  return reinterpret_cast<cpx<T>*>(p->w1);
}
template <typename T>
cpx<T>* W2(const fsx_plan_s* p) {
  return reinterpret_cast<cpx<T>*>(p->w2);
}

size_t kelems(const fsx_plan_s* p) { return (size_t)p->nz * p->ny * p->nkx; }
size_t xelems(const fsx_plan_s* p) { return (size_t)p->nz * p->ny * p->nx; }
int nrows(const fsx_plan_s* p) { return p->nz * p->ny; }

// ---------------------------------------------------------------- inverse
// K (nf fields, field stride kfs) -> W2 through the z and y passes.
template <typename T>
int inverse_to_w2(fsx_plan_s* p, const cpx<T>* kin, long long kfs, int nf,
                  bool pruned, hipStream_t s) {
  const Geo g = geo<T>(p, pruned);
  if (p->dim == 2) {
    OpYInv2<T> op;
    op.in = kin;
    op.out = W2<T>(p);
    op.ifs = kfs;
    op.ofs = g.fs2;
    op.nkx = p->nkx;
    op.ld2 = g.ld2;
    op.ym = g.ym;
    note(K_YINV, (double)nf * g.mx * (g.ym.nk() + p->ny) * sizeof(cpx<T>));
    return launch_strided<T, +1>(op, p->ny, 1, g.mx, nf, p->tw[1], s, true);
  }
  {
    OpZInv<T> op;
    op.in = kin;
    op.out = W1<T>(p);
    op.ifs = kfs;
    op.ofs = g.fs1;
    op.ny = p->ny;
    op.nkx = p->nkx;
    op.ld1 = g.ld1;
    op.ym = g.ym;
    op.zm = g.zm;
    note(K_ZINV, (double)nf * g.ym.nk() * g.mx * (g.zm.nk() + p->nz) * sizeof(cpx<T>));
    int rc = launch_strided<T, +1>(op, p->nz, g.ym.nk(), g.mx, nf, p->tw[2], s, true);
    if (rc) return rc;
  }
  OpYInv3<T> op;
  op.in = W1<T>(p);
  op.out = W2<T>(p);
  op.ifs = g.fs1;
  op.ofs = g.fs2;
  op.ny = p->ny;
  op.ld1 = g.ld1;
  op.ld2 = g.ld2;
  op.ym = g.ym;
  note(K_YINV, (double)nf * p->nz * g.mx * (g.ym.nk() + p->ny) * sizeof(cpx<T>));
  return launch_strided<T, +1>(op, p->ny, p->nz, g.mx, nf, p->tw[1], s);
}

// W2 (nf fields) -> K through the y and z passes; only kept modes written.
template <typename T>
int forward_from_w2(fsx_plan_s* p, cpx<T>* kout, long long kfs, int nf,
                    bool pruned, hipStream_t s) {
  const Geo g = geo<T>(p, pruned);
  const unsigned char* mask = pruned ? p->d_mask : nullptr;
  if (p->dim == 2) {
    OpYFwd2<T> op;
    op.in = W2<T>(p);
    op.out = kout;
    op.ifs = g.fs2;
    op.ofs = kfs;
    op.nkx = p->nkx;
    op.ld2 = g.ld2;
    op.ym = g.ym;
    op.mask = mask;
    note(K_YFWD, (double)nf * g.mx * (g.ym.nk() + p->ny) * sizeof(cpx<T>));
    return launch_strided<T, -1>(op, p->ny, 1, g.mx, nf, p->tw[1], s, true);
  }
  {
    OpYFwd3<T> op;
    op.in = W2<T>(p);
    op.out = W1<T>(p);
    op.ifs = g.fs2;
    op.ofs = g.fs1;
    op.ny = p->ny;
    op.ld1 = g.ld1;
    op.ld2 = g.ld2;
    op.ym = g.ym;
    note(K_YFWD, (double)nf * p->nz * g.mx * (g.ym.nk() + p->ny) * sizeof(cpx<T>));
    int rc = launch_strided<T, -1>(op, p->ny, p->nz, g.mx, nf, p->tw[1], s);
    if (rc) return rc;
  }
  OpZFwd<T> op;
  op.in = W1<T>(p);
  op.out = kout;
  op.ifs = g.fs1;
  op.ofs = kfs;
  op.ny = p->ny;
  op.nkx = p->nkx;
  op.ld1 = g.ld1;
  op.ym = g.ym;
  op.zm = g.zm;
  op.mask = mask;
  note(K_ZFWD, (double)nf * g.ym.nk() * g.mx * (p->nz + g.zm.nk()) * sizeof(cpx<T>));
  return launch_strided<T, -1>(op, p->nz, g.ym.nk(), g.mx, nf, p->tw[2], s, true);
}

template <typename T>
int do_ifft(fsx_plan_s* p, const void* kin, void* xout, int nfields,
            hipStream_t s) {
  const Geo g = geo<T>(p, false);
  for (int f = 0; f < nfields; ++f) {
    const cpx<T>* k = reinterpret_cast<const cpx<T>*>(kin) + f * kelems(p);
    T* x = reinterpret_cast<T*>(xout) + f * xelems(p);
    int rc = inverse_to_w2<T>(p, k, 0, 1, false, s);
    if (rc) return rc;
    OpXC2R<T> op;
    op.in = W2<T>(p);
    op.outr = x;
    note(K_XPASS, (double)nrows(p) * (p->mxF * sizeof(cpx<T>) + p->nx * sizeof(T)));
    rc = launch_xpass<T>(op, p->nx, nrows(p), 1, 1, p->mxF, g.ld2, p->mxF,
                         g.ld2, 0, 0, T(1), p->tw[0], s);
    if (rc) return rc;
  }
  return 0;
}

template <typename T>
int do_fft(fsx_plan_s* p, const void* xin, void* kout, int nfields,
           hipStream_t s) {
  const Geo g = geo<T>(p, false);
  const double npts = (double)p->nx * p->ny * p->nz;
  for (int f = 0; f < nfields; ++f) {
    const T* x = reinterpret_cast<const T*>(xin) + f * xelems(p);
    cpx<T>* k = reinterpret_cast<cpx<T>*>(kout) + f * kelems(p);
    OpXR2C<T> op;
    op.inr = x;
    op.out = W2<T>(p);
    note(K_XPASS, (double)nrows(p) * (p->mxF * sizeof(cpx<T>) + p->nx * sizeof(T)));
    int rc = launch_xpass<T>(op, p->nx, nrows(p), 1, 1, p->mxF, g.ld2, p->mxF,
                             g.ld2, 0, 0, T(1.0 / npts), p->tw[0], s);
    if (rc) return rc;
    rc = forward_from_w2<T>(p, k, 0, 1, false, s);
    if (rc) return rc;
  }
  return 0;
}

// ns2d: y-pass inverse (u, v, grad rot) and row pass (products, r2c) into
// W2 field 0.
template <typename T>
int ns2d_front(fsx_plan_s* p, const void* rot, double beta, int want_maxima,
               hipStream_t s) {
  const Geo g = geo<T>(p, true);
  OpYInv2NS<T> op;
  op.in = reinterpret_cast<const cpx<T>*>(rot);
  op.out = W2<T>(p);
  op.ofs = g.fs2;
  op.ny = p->ny;
  op.nkx = p->nkx;
  op.ld2 = g.ld2;
  op.ym = g.ym;
  op.dkx = (T)p->dkx;
  op.dky = (T)p->dky;
  note(K_YINV, (double)g.mx * (g.ym.nk() + 4.0 * p->ny) * sizeof(cpx<T>));
  int rc;
  if (p->ysplit) {
    OpYInv2NSOne<T> op1;
    static_cast<OpYInv2NS<T>&>(op1) = op;
    rc = launch_strided<T, +1>(op1, p->ny, 1, g.mx, 4, p->tw[1], s, true, p->ylines_inv);
  } else {
    rc = launch_strided<T, +1>(op, p->ny, 1, g.mx, 4, p->tw[1], s, true, p->ylines_inv);
  }
  if (rc) return rc;
  unsigned long long* mxs =
      want_maxima ? reinterpret_cast<unsigned long long*>(p->d_scal) : nullptr;
  const double npts = (double)p->nx * p->ny;
  note(K_XPASS, 5.0 * p->ny * g.mx * sizeof(cpx<T>));
  XRNS2D<T> xr;
  xr.beta = (T)beta;
  rc = launch_xrow<T>(p->nx, p->ny, g.mx, g.ld2, g.mx, g.ld2, g.fs2, g.fs2,
                      T(1.0 / npts), W2<T>(p), W2<T>(p), p->tw[0], mxs, xr, s);
  if (rc != 1) return rc;
  OpXNS2D<T, kXThr> ox;
  ox.in = W2<T>(p);
  ox.out = W2<T>(p);
  ox.maxima = mxs;
  ox.beta = (T)beta;
  return launch_xpass<T>(ox, p->nx, p->ny, 4, 1, g.mx, g.ld2, g.mx, g.ld2,
                         g.fs2, g.fs2, T(1.0 / npts), p->tw[0], s);
}

template <typename T>
int do_ns2d(fsx_plan_s* p, const void* rot, void* out, double beta,
            int want_maxima, hipStream_t s) {
  int rc = ns2d_front<T>(p, rot, beta, want_maxima, s);
  if (rc) return rc;
  return forward_from_w2<T>(p, reinterpret_cast<cpx<T>*>(out), 0, 1, true, s);
}

// ns3d (strat = 0): z-pass with curl, y-pass, row pass (v x omega, r2c)
// into W2 fields 0-2.  ns3d.strat (strat = 1): v, omega and b through the
// z- and y-passes, row pass producing v x omega and v b into W2 fields 0-5.
template <typename T>
int ns3d_front(fsx_plan_s* p, const void* v, int has_f, double fcor,
               int want_maxima, hipStream_t s, int strat = 0) {
  const Geo g = geo<T>(p, true);
  const long long ke = (long long)kelems(p);
  const int nin = strat ? 7 : 6, nout = strat ? 6 : 3;
  OpZInvCurl<T> op;
  op.in = reinterpret_cast<const cpx<T>*>(v);
  op.out = W1<T>(p);
  op.ifs = ke;
  op.ofs = g.fs1;
  op.ny = p->ny;
  op.nz = p->nz;
  op.nkx = p->nkx;
  op.ld1 = g.ld1;
  op.ym = g.ym;
  op.zm = g.zm;
  op.dkx = (T)p->dkx;
  op.dky = (T)p->dky;
  op.dkz = (T)p->dkz;
  op.has_f = has_f;
  op.fcor = (T)fcor;
  op.with_b = strat;
  note(K_ZINV, (double)g.ym.nk() * g.mx *
                   ((nin - 3.0) * g.zm.nk() + (double)nin * p->nz) * sizeof(cpx<T>));
  int rc = launch_strided<T, +1>(op, p->nz, g.ym.nk(), g.mx, nin, p->tw[2], s, true);
  if (rc) return rc;
  OpYInv3<T> oy;
  oy.in = W1<T>(p);
  oy.out = W2<T>(p);
  oy.ifs = g.fs1;
  oy.ofs = g.fs2;
  oy.ny = p->ny;
  oy.ld1 = g.ld1;
  oy.ld2 = g.ld2;
  oy.ym = g.ym;
  note(K_YINV, (double)nin * p->nz * g.mx * (g.ym.nk() + p->ny) * sizeof(cpx<T>));
  rc = launch_strided<T, +1>(oy, p->ny, p->nz, g.mx, nin, p->tw[1], s);
  if (rc) return rc;
  unsigned long long* mx =
      want_maxima ? reinterpret_cast<unsigned long long*>(p->d_scal) : nullptr;
  const double npts = (double)p->nx * p->ny * p->nz;
  note(K_XPASS, (double)(nin + nout) * nrows(p) * g.mx * sizeof(cpx<T>));
  if (strat) {
    XRStrat<T> xr;
    rc = launch_xrow<T>(p->nx, nrows(p), g.mx, g.ld2, g.mx, g.ld2, g.fs2, g.fs2,
                        T(1.0 / npts), W2<T>(p), W2<T>(p), p->tw[0], mx, xr, s);
    if (rc != 1) return rc;
  } else {
    XRNS3D<T> xr;
    rc = launch_xrow<T>(p->nx, nrows(p), g.mx, g.ld2, g.mx, g.ld2, g.fs2, g.fs2,
                        T(1.0 / npts), W2<T>(p), W2<T>(p), p->tw[0], mx, xr, s);
    if (rc != 1) return rc;
  }
  if (strat) {
    OpXStrat<T, kXThr> ox;
    ox.in = W2<T>(p);
    ox.out = W2<T>(p);
    ox.maxima = mx;
    return launch_xpass<T>(ox, p->nx, nrows(p), nin, nout, g.mx, g.ld2, g.mx,
                           g.ld2, g.fs2, g.fs2, T(1.0 / npts), p->tw[0], s);
  }
  OpXNS3D<T, kXThr> ox;
  ox.in = W2<T>(p);
  ox.out = W2<T>(p);
  ox.maxima = mx;
  return launch_xpass<T>(ox, p->nx, nrows(p), nin, nout, g.mx, g.ld2, g.mx,
                         g.ld2, g.fs2, g.fs2, T(1.0 / npts), p->tw[0], s);
}

template <typename T>
int do_ns3d(fsx_plan_s* p, const void* v, void* out, int has_f, double fcor,
            int want_maxima, hipStream_t s) {
  int rc = ns3d_front<T>(p, v, has_f, fcor, want_maxima, s);
  if (rc) return rc;
  return forward_from_w2<T>(p, reinterpret_cast<cpx<T>*>(out),
                            (long long)kelems(p), 3, true, s);
}

template <typename T>
RKArgs<T> make_rk(fsx_plan_s* p, int scheme, int stage, int nvar, void* S,
                  void* tmp, void* t1, void* N, int project, int no_vz_kz0,
                  const double* visc) {
  RKArgs<T> a;
  const size_t ke = kelems(p);
  for (int v = 0; v < 4; ++v) {
    const bool on = v < nvar;
    a.S[v] = on ? reinterpret_cast<cpx<T>*>(S) + v * ke : nullptr;
    a.tmp[v] = on && tmp ? reinterpret_cast<cpx<T>*>(tmp) + v * ke : nullptr;
    a.t1[v] = on ? reinterpret_cast<cpx<T>*>(t1) + v * ke : nullptr;
    a.Nt[v] = on && N ? reinterpret_cast<cpx<T>*>(N) + v * ke : nullptr;
  }
  a.nvar = nvar;
  a.scheme = scheme;
  a.stage = stage;
  a.project = project;
  a.no_vz_kz0 = no_vz_kz0;
  a.dt = p->d_scal + 4;
  a.visc.nu2 = visc[0];
  a.visc.nu4 = visc[1];
  a.visc.nu8 = visc[2];
  a.visc.num4 = visc[3];
  a.box = make_box(p, true);
  a.nanflag = reinterpret_cast<unsigned int*>(p->d_scal + 5);
  // the Runge-Kutta stages add the forcing; tendency-only calls do not
  const bool forced = p->force_on && p->d_fc && stage >= 0;
  a.fc = forced ? reinterpret_cast<const cpx<T>*>(p->d_fc) : nullptr;
  a.fvs = p->fvs;
  a.fnyc = p->fnyc;
  a.fnzc = p->fnzc;
  a.fnkxc = p->fnkxc;
  for (int v = 0; v < 4; ++v) a.ext[v] = nullptr;
  return a;
}

// Normalised forcing of the step (k_forcing_norm) on the state S.
template <typename T>
int forcing_normalize(fsx_plan_s* p, const void* S, hipStream_t s) {
  if (!p->d_fc || !p->d_fmodes || p->fnmodes <= 0)
    return fail(FSX_ERR_ARG, "forcing box not set up (fsx_forcing_setup)");
  ForcingArgs<T> f;
  f.S = reinterpret_cast<const cpx<T>*>(S);
  f.ke = (long long)kelems(p);
  f.fc = reinterpret_cast<cpx<T>*>(p->d_fc);
  f.fvs = p->fvs;
  f.nvar = p->fnvar;
  f.modes = p->d_fmodes;
  f.w = p->d_fw;
  f.f0 = reinterpret_cast<const cpx<double>*>(p->d_ff0);
  f.f1 = reinterpret_cast<const cpx<double>*>(p->d_ff1);
  f.nmodes = p->fnmodes;
  f.nyc = p->fnyc;
  f.nzc = p->fnzc;
  f.nkxc = p->fnkxc;
  f.ny = p->ny;
  f.nz = p->dim == 3 ? p->nz : 1;
  f.nkx = p->nkx;
  f.dkx = p->dkx;
  f.dky = p->dky;
  f.key_kind = p->fkey_kind;
  f.key_var = p->fkey_var;
  f.coef = p->fcoef;
  f.rate = p->frate;
  f.which_root = p->froot;
  f.dt = p->d_scal + 4;
  f.par = p->fgraph ? p->d_scal + 16 : nullptr;
  hipLaunchKernelGGL(k_forcing_norm<T>, dim3(1), dim3(256), 0, s, f);
  FSX_HIP(hipGetLastError());
  return 0;
}

// Stage 0 of a fused step: the pending normalisation runs after the CFL
// update (the dt of the step) and before the stage algebra, on S.
template <typename T>
int forcing_stage0(fsx_plan_s* p, int stage, const void* S, hipStream_t s) {
  if (stage != 0 || !p->force_pending) return 0;
  p->force_pending = false;
  return forcing_normalize<T>(p, S, s);
}

// algorithmic bytes of the state streams of one Runge-Kutta stage
double rk_bytes(const fsx_plan_s* p, int scheme, int stage, int nvar, int csz) {
  const double kept = (double)(p->dim == 3 ? p->zmP.nk() : 1) * p->ymP.nk() * p->mxP;
  const int streams = scheme == 4 ? (stage == 0 ? 3 : (stage == 3 ? 2 : 4))
                                  : (stage == 0 ? 2 : 2);
  return kept * nvar * streams * csz;
}

template <typename T>
int do_rk(fsx_plan_s* p, int scheme, int stage, int nvar, void* S, void* tmp,
          void* t1, void* N, int project, int no_vz_kz0, const double* visc,
          hipStream_t s) {
  RKArgs<T> a = make_rk<T>(p, scheme, stage, nvar, S, tmp, t1, N, project,
                           no_vz_kz0, visc);
  const double kept = (double)(p->dim == 3 ? p->zmP.nk() : 1) * p->ymP.nk() * p->mxP;
  note(K_RK, rk_bytes(p, scheme, stage, nvar, sizeof(cpx<T>)) +
                 kept * nvar * sizeof(cpx<T>));
  int rc = tm_pre(s);
  if (rc) return rc;
  const int nrow = (p->dim == 3 ? p->zmP.nk() : 1) * p->ymP.nk();
  unsigned blocks = (unsigned)((nrow + 3) / 4);
  if (blocks > 16384) blocks = 16384;
  if (nvar == 1)
    hipLaunchKernelGGL((k_rk_stage<T, 1>), dim3(blocks), dim3(256), 0, s, a);
  else if (nvar == 2)
    hipLaunchKernelGGL((k_rk_stage<T, 2>), dim3(blocks), dim3(256), 0, s, a);
  else if (nvar == 3)
    hipLaunchKernelGGL((k_rk_stage<T, 3>), dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_rk_stage<T, 4>), dim3(blocks), dim3(256), 0, s, a);
  FSX_HIP(hipGetLastError());
  return tm_post(s);
}

int cfl_begin(fsx_plan_s* p, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s,
                     reinterpret_cast<unsigned long long*>(p->d_scal), 4, 0ull);
  FSX_HIP(hipGetLastError());
  return 0;
}

int cfl_end(fsx_plan_s* p, int ncomp, double cfl, double dt_max, hipStream_t s) {
  const double dx = p->Lx / p->nx, dy = p->Ly / p->ny,
               dz = p->dim == 3 ? p->Lz / p->nz : 1.0;
  hipLaunchKernelGGL(k_cfl_update, dim3(1), dim3(64), 0, s,
                     reinterpret_cast<const unsigned long long*>(p->d_scal),
                     ncomp, dx, dy, dz, cfl, dt_max, p->d_scal + 4);
  FSX_HIP(hipGetLastError());
  return 0;
}

// Fused ns2d Runge-Kutta stage: tendency of S (stage 0) or t1, optional CFL
// update of the device dt at stage 0, forward y-pass with the stage algebra.
template <typename T>
int do_ns2d_rk(fsx_plan_s* p, int scheme, int stage, void* S, void* tmp,
               void* t1, double beta, const double* visc, double cfl,
               double dt_max, hipStream_t s, const void* ext = nullptr) {
  const bool use_cfl = cfl > 0 && stage == 0;
  int rc = 0;
  if (use_cfl && (rc = cfl_begin(p, s))) return rc;
  rc = ns2d_front<T>(p, stage == 0 ? S : t1, beta, use_cfl, s);
  if (rc) return rc;
  if (use_cfl && (rc = cfl_end(p, 2, cfl, dt_max, s))) return rc;
  if ((rc = forcing_stage0<T>(p, stage, S, s))) return rc;
  const Geo g = geo<T>(p, true);
  OpYFwd2RK<T> op;
  op.in = W2<T>(p);
  op.ny = p->ny;
  op.nkx = p->nkx;
  op.ld2 = g.ld2;
  op.ym = g.ym;
  op.mask = p->d_mask;
  op.rk = make_rk<T>(p, scheme, stage, 1, S, tmp, t1, nullptr, 0, 0, visc);
  op.rk.ext[0] = reinterpret_cast<const cpx<T>*>(ext);
  note(K_ZFWD, (double)g.mx * p->ny * sizeof(cpx<T>) +
                   rk_bytes(p, scheme, stage, 1, sizeof(cpx<T>)));
  return launch_strided<T, -1>(op, p->ny, 1, g.mx, 1, p->tw[1], s, true, p->ylines_fwd);
}

// Batched ns2d Runge-Kutta stage over the nb runs of a batch plan: the
// same three passes as do_ns2d_rk with the run as the outer index of the
// y-passes (nouter = nb) and nb * ny rows in the row pass; each run's dt
// and viscosities come from d_bpar.  Fixed dt (no CFL), no forcing.
template <typename T>
int do_ns2d_batch_rk(fsx_plan_s* p, int scheme, int stage, void* S, void* tmp, void* t1,
                     double beta, hipStream_t s) {
  const Geo g = geo<T>(p, true);
  const int nb = p->nb;
  const long long ke = (long long)kelems(p);      // modes of one run
  const long long fs2 = (long long)nb * p->ny * g.ld2;  // W2 field stride
  OpYInv2NS<T> op;
  op.in = reinterpret_cast<const cpx<T>*>(stage == 0 ? S : t1);
  op.out = W2<T>(p);
  op.ofs = fs2;
  op.ips = ke;
  op.ops = (long long)p->ny * g.ld2;
  op.ny = p->ny;
  op.nkx = p->nkx;
  op.ld2 = g.ld2;
  op.ym = g.ym;
  op.dkx = (T)p->dkx;
  op.dky = (T)p->dky;
  note(K_YINV, (double)nb * g.mx * (g.ym.nk() + 4.0 * p->ny) * sizeof(cpx<T>));
  int rc = launch_strided<T, +1>(op, p->ny, nb, g.mx, 4, p->tw[1], s, true, p->blines);
  if (rc) return rc;
  const int rows = nb * p->ny;
  const double npts = (double)p->nx * p->ny;
  note(K_XPASS, 5.0 * rows * g.mx * sizeof(cpx<T>));
  XRNS2D<T> xr;
  xr.beta = (T)beta;
  // row pass of the batch (fsx_batch_set_row): 0 register rows (8 points
  // per lane in FP32), 1 register rows with 4 points per lane, 2 the LDS
  // row pass
  rc = p->brow == 2 ? 1
                    : launch_xrow<T>(p->nx, rows, g.mx, g.ld2, g.mx, g.ld2, fs2, fs2,
                                     T(1.0 / npts), W2<T>(p), W2<T>(p), p->tw[0], nullptr, xr,
                                     s, p->brow == 1);
  if (rc == 1) {
    OpXNS2D<T, kXThr> ox;
    ox.in = W2<T>(p);
    ox.out = W2<T>(p);
    ox.maxima = nullptr;
    ox.beta = (T)beta;
    rc = launch_xpass<T>(ox, p->nx, rows, 4, 1, g.mx, g.ld2, g.mx, g.ld2, fs2, fs2,
                         T(1.0 / npts), p->tw[0], s);
  }
  if (rc) return rc;
  OpYFwd2RK<T> ob;
  ob.in = W2<T>(p);
  ob.ny = p->ny;
  ob.nkx = p->nkx;
  ob.ld2 = g.ld2;
  ob.ym = g.ym;
  ob.mask = p->d_mask;
  ob.ips = (long long)p->ny * g.ld2;
  ob.ops = ke;
  ob.bpar = p->d_bpar;
  const double visc0[4] = {0, 0, 0, 0};
  // state arrays hold the nb runs one after the other (one variable)
  ob.rk = make_rk<T>(p, scheme, stage, 1, S, tmp, t1, nullptr, 0, 0, visc0);
  ob.rk.fc = nullptr;
  note(K_ZFWD, (double)nb * (g.mx * p->ny * sizeof(cpx<T>) +
                             rk_bytes(p, scheme, stage, 1, sizeof(cpx<T>))));
  return launch_strided<T, -1>(ob, p->ny, nb, g.mx, 1, p->tw[1], s, true, p->blines);
}

template <typename T>
int do_ns3d_rk(fsx_plan_s* p, int scheme, int stage, void* S, void* tmp,
               void* t1, int has_f, double fcor, int project, int no_vz_kz0,
               const double* visc, double cfl, double dt_max, hipStream_t s) {
  const bool use_cfl = cfl > 0 && stage == 0;
  int rc = 0;
  if (use_cfl && (rc = cfl_begin(p, s))) return rc;
  rc = ns3d_front<T>(p, stage == 0 ? S : t1, has_f, fcor, use_cfl, s);
  if (rc) return rc;
  const Geo g = geo<T>(p, true);
  {
    OpYFwd3<T> op;
    op.in = W2<T>(p);
    op.out = W1<T>(p);
    op.ifs = g.fs2;
    op.ofs = g.fs1;
    op.ny = p->ny;
    op.ld1 = g.ld1;
    op.ld2 = g.ld2;
    op.ym = g.ym;
    note(K_YFWD, 3.0 * p->nz * g.mx * (g.ym.nk() + p->ny) * sizeof(cpx<T>));
    rc = launch_strided<T, -1>(op, p->ny, p->nz, g.mx, 3, p->tw[1], s);
    if (rc) return rc;
  }
  if (use_cfl && (rc = cfl_end(p, 3, cfl, dt_max, s))) return rc;
  if ((rc = forcing_stage0<T>(p, stage, S, s))) return rc;
  OpZFwdRK3<T> op;
  op.in = W1<T>(p);
  op.ifs = g.fs1;
  op.ny = p->ny;
  op.nkx = p->nkx;
  op.ld1 = g.ld1;
  op.ym = g.ym;
  op.zm = g.zm;
  op.mask = p->d_mask;
  op.rk = make_rk<T>(p, scheme, stage, 3, S, tmp, t1, nullptr, project, no_vz_kz0,
                     visc);
  note(K_ZFWD, 3.0 * g.ym.nk() * g.mx * p->nz * sizeof(cpx<T>) +
                   rk_bytes(p, scheme, stage, 3, sizeof(cpx<T>)));
  return launch_strided<T, -1>(op, p->nz, g.ym.nk(), g.mx, 3, p->tw[2], s, true);
}

template <typename T>
int do_sums(fsx_plan_s* p, int kind, const void* f0, const void* f1,
            const void* f2, const void* f3, double N2, const double* visc,
            double* host_out, hipStream_t s) {
  Visc v = {visc[0], visc[1], visc[2], visc[3]};
  double* d = p->d_scal + 8;
  FSX_HIP(hipMemsetAsync(d, 0, 6 * sizeof(double), s));
  const size_t n = kelems(p);
  hipLaunchKernelGGL(k_sums<T>, dim3(grid_for(n)), dim3(256), 0, s,
                     reinterpret_cast<const cpx<T>*>(f0),
                     reinterpret_cast<const cpx<T>*>(f1),
                     reinterpret_cast<const cpx<T>*>(f2),
                     reinterpret_cast<const cpx<T>*>(f3), kind, N2,
                     make_box(p, false), v, d);
  FSX_HIP(hipGetLastError());
  FSX_HIP(hipMemcpyAsync(host_out, d, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
  FSX_HIP(hipStreamSynchronize(s));
  return 0;
}

// ns3d.strat: one fused Runge-Kutta stage (stage >= 0) or the drop-in
// tendency into N (stage = -1, with projection and dealiasing).
template <typename T>
int do_strat(fsx_plan_s* p, int scheme, int stage, void* S, void* tmp, void* t1,
             void* N, const void* src, int has_f, double fcor, double bvf2,
             int project, int no_vz_kz0, const double* visc, double cfl,
             double dt_max, bool want_max, hipStream_t s) {
  const bool use_cfl = cfl > 0 && stage == 0;
  const bool reduce = use_cfl || want_max;
  int rc = 0;
  if (reduce && (rc = cfl_begin(p, s))) return rc;
  rc = ns3d_front<T>(p, src, has_f, fcor, reduce, s, 1);
  if (rc) return rc;
  const Geo g = geo<T>(p, true);
  {
    OpYFwd3<T> op;
    op.in = W2<T>(p);
    op.out = W1<T>(p);
    op.ifs = g.fs2;
    op.ofs = g.fs1;
    op.ny = p->ny;
    op.ld1 = g.ld1;
    op.ld2 = g.ld2;
    op.ym = g.ym;
    note(K_YFWD, 6.0 * p->nz * g.mx * (g.ym.nk() + p->ny) * sizeof(cpx<T>));
    rc = launch_strided<T, -1>(op, p->ny, p->nz, g.mx, 6, p->tw[1], s);
    if (rc) return rc;
  }
  if (use_cfl && (rc = cfl_end(p, 3, cfl, dt_max, s))) return rc;
  if ((rc = forcing_stage0<T>(p, stage, S, s))) return rc;
  OpZFwdRKStrat<T> op;
  op.in = W1<T>(p);
  op.ifs = g.fs1;
  op.ny = p->ny;
  op.nkx = p->nkx;
  op.ld1 = g.ld1;
  op.ym = g.ym;
  op.zm = g.zm;
  op.mask = p->d_mask;
  op.dkx = (T)p->dkx;
  op.dky = (T)p->dky;
  op.dkz = (T)p->dkz;
  op.N2 = (T)bvf2;
  op.rk = make_rk<T>(p, scheme, stage, 4, S, tmp, t1, N, project, no_vz_kz0, visc);
  const size_t ke = kelems(p);
  for (int v = 0; v < 4; ++v)
    op.rk.src[v] = reinterpret_cast<const cpx<T>*>(src) + v * ke;
  const double kept = (double)p->zmP.nk() * p->ymP.nk() * p->mxP;
  const double rkb = stage < 0 ? kept * 4 * sizeof(cpx<T>)
                               : rk_bytes(p, scheme, stage, 4, sizeof(cpx<T>));
  note(K_ZFWD, 6.0 * g.ym.nk() * g.mx * p->nz * sizeof(cpx<T>) +
                   2.0 * kept * sizeof(cpx<T>) + rkb);
  return launch_strided<T, -1>(op, p->nz, g.ym.nk(), g.mx, 6, p->tw[2], s, true);
}

unsigned grid_blocks(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

KGrid kgrid(const fsx_plan_s* p) {
  KGrid g;
  g.nz = p->nz;
  g.ny = p->ny;
  g.nkx = p->nkx;
  g.dkx = p->dkx;
  g.dky = p->dky;
  g.dkz = p->dkz;
  return g;
}

}  // namespace
