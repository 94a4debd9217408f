// Plans, launches and the C-ABI of the fluidsim_amd engine (gfx950).
//
// The exported functions are declared in include/fluidsim_amd.h.  Every
// function returns 0 on success or a negative code; fsx_last_error() gives
// the message.  Device buffers are plain pointers (complex = interleaved
// re/im of the plan precision), streams are hipStream_t passed as void*.

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

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
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
  double* d_scal;  // [0..3] maxima bits, [4] dt, [5] nan flag, [8..15] sums
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
};

namespace {

// kinds of launches reported by fsx_timing_read
enum { K_ZINV = 0, K_YINV = 1, K_XPASS = 2, K_YFWD = 3, K_ZFWD = 4, K_RK = 5,
       K_OTHER = 6, K_DIAG = 7, K_CNN = 8 };

thread_local fsx_plan_s* g_plan = nullptr;  // plan of the running API call
thread_local int g_kind = K_OTHER;
thread_local double g_bytes = 0;

struct PlanScope {
  explicit PlanScope(fsx_plan_s* p) { g_plan = p; }
  ~PlanScope() { g_plan = nullptr; }
};

void note(int kind, double bytes) {
  g_kind = kind;
  g_bytes = bytes;
}

int tm_flush(fsx_plan_s* p);

int tm_pre(hipStream_t s) {
  fsx_plan_s* p = g_plan;
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
  fsx_plan_s* p = g_plan;
  if (!p || !p->tm || !p->tm->on) return 0;
  fsx_timing* t = p->tm;
  FSX_HIP(hipEventRecord(t->ev[2 * t->used + 1], s));
  t->kind[t->used] = g_kind;
  t->bytes[t->used] = g_bytes;
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
#ifndef FSX_XPASS_THREADS
#define FSX_XPASS_THREADS 512
#endif
#ifndef FSX_XPASS_EPT_DIV
#define FSX_XPASS_EPT_DIV 2
#endif
constexpr int kXThr = FSX_XPASS_THREADS;  // threads per workgroup of the row pass
constexpr int kEPT32 = 16;  // 8192 complex64 per tile  = 64 KiB of LDS
constexpr int kEPT64 = 8;   // 4096 complex128 per tile = 64 KiB of LDS
#ifndef FSX_MIN_BLOCKS
#define FSX_MIN_BLOCKS 1024
#endif
constexpr int kMinBlocks = FSX_MIN_BLOCKS;  // shrink tiles of small problems below this

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
template <typename T, int DIR, int E, int KT, class Op>
int launch_strided_tile(Op& op, int N, size_t tile_max, long long blocks, int nfb,
                        hipStream_t s) {
  const size_t tile = std::min(tile_max, (size_t)E * KT);
  const size_t lds = (tile + tile / 32 + N) * sizeof(cpx<T>);
  auto kern = k_strided<T, DIR, E, KT, Op>;
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
  const bool wide_ok = sizeof(T) == 4 && tile / N < wide_below &&
                       ((size_t)E * KW + E * KW / 32 + N) * sizeof(cpx<T>) <= kLdsMax;
  int L = (wide_ok ? E * KW : tile) / N;
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
  if (!wide && need * line_mul<Op>::value <= (size_t)E * kThr) need *= line_mul<Op>::value;
  if (wide) return launch_strided_tile<T, DIR, E, KW>(op, N, need, blocks, nfb, s);
  // small problems (few lines per tile): as many threads as the tile needs
  if ((size_t)E * kThr / 4 >= need)
    return launch_strided_tile<T, DIR, E, kThr / 4>(op, N, need, blocks, nfb, s);
  if ((size_t)E * kThr / 2 >= need)
    return launch_strided_tile<T, DIR, E, kThr / 2>(op, N, need, blocks, nfb, s);
  return launch_strided_tile<T, DIR, E, kThr>(op, N, need, blocks, nfb, s);
}

template <typename T, class Op>
int launch_xpass(Op& op, int nx, int nrows, int nf_in, int nf_out, int mx_in,
                 int ld_in, int mx_out, int ld_out, long long ifs,
                 long long ofs, T scl, const void* tw, hipStream_t s) {
  constexpr int E = ept<T>() / FSX_XPASS_EPT_DIV >= 8 ? ept<T>() / FSX_XPASS_EPT_DIV : 8;
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
#ifndef FSX_XROW
#define FSX_XROW 1
#endif
#ifndef FSX_XROW_THREADS
#define FSX_XROW_THREADS 256
#endif
#ifndef FSX_XROW_BLOCKS_PER_CU
#define FSX_XROW_BLOCKS_PER_CU 4
#endif
constexpr int kXRThr = FSX_XROW_THREADS;

#ifndef FSX_XR_LOGP32
#define FSX_XR_LOGP32 3
#endif
#ifndef FSX_XR_LOGP64
#define FSX_XR_LOGP64 2
#endif
template <typename T>
constexpr int xrow_logp() {
  return sizeof(T) == 4 ? FSX_XR_LOGP32 : FSX_XR_LOGP64;
}

template <typename T, int LOGM, class Op>
int launch_xrow_m(const XRDesc<T>& d, const Op& op, hipStream_t s) {
  constexpr int LOGP = xrow_logp<T>();
  using PL = RowPlan<LOGM, LOGP>;
  constexpr int RPB = (kXRThr / 64) * (64 / PL::G);
  const size_t lds = ((size_t)RPB * xrow_ld<T>(PL::M) + (FSX_XR_TWLDS ? 2 * PL::M : 0)) *
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
  const long long cap = (long long)ncu * FSX_XROW_BLOCKS_PER_CU;
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
                hipStream_t s) {
  if (!FSX_XROW) return 1;
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
  int rc = launch_strided<T, +1>(op, p->ny, 1, g.mx, 4, p->tw[1], s, true);
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
  return launch_strided<T, -1>(op, p->ny, 1, g.mx, 1, p->tw[1], s, true);
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
  rc = launch_xrow<T>(p->nx, rows, g.mx, g.ld2, g.mx, g.ld2, fs2, fs2, T(1.0 / npts),
                      W2<T>(p), W2<T>(p), p->tw[0], nullptr, xr, s);
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

}  // namespace

// ====================================================================== ABI
extern "C" {

int fsx_version(void) { return FSX_ABI_VERSION; }

int fsx_last_error(char* buf, int len) {
  if (!buf || len <= 0) return -1;
  std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
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
  if (!is_pow2(nx) || !is_pow2(ny) || !is_pow2(nz) || nx < 4 || ny < 2 ||
      (dim == 3 && nz < 2))
    return fail(FSX_ERR_ARG,
                "the engine supports power-of-two grids (nx >= 4, ny, nz >= 2)");
  const int tile = (precision == 4 ? kEPT32 : kEPT64) * kThr;
  if (ny > tile || nz > tile || nx / 2 > tile)
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

int fsx_ns2d_tendency(fsx_plan p, const void* rot_fft, void* out_fft,
                      double beta, int want_maxima, void* stream) {
  PlanScope scope_(p);
  if (!p || !rot_fft || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2) return fail(FSX_ERR_ARG, "ns2d needs a 2D plan");
  if (p->nf_max < 4) return fail(FSX_ERR_ARG, "ns2d needs nf_max >= 4");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d<float>(p, rot_fft, out_fft, beta, want_maxima, s)
                      : do_ns2d<double>(p, rot_fft, out_fft, beta, want_maxima, s);
}

int fsx_ns3d_tendency(fsx_plan p, const void* v_fft, void* out_fft, int has_f,
                      double f, int want_maxima, void* stream) {
  PlanScope scope_(p);
  if (!p || !v_fft || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3) return fail(FSX_ERR_ARG, "ns3d needs a 3D plan");
  if (p->nf_max < 6) return fail(FSX_ERR_ARG, "ns3d needs nf_max >= 6");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4
             ? do_ns3d<float>(p, v_fft, out_fft, has_f, f, want_maxima, s)
             : do_ns3d<double>(p, v_fft, out_fft, has_f, f, want_maxima, s);
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

int fsx_ns2d_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                      void* t1, double beta, const double* visc, double cfl,
                      double dt_max, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2 || p->nf_max < 4) return fail(FSX_ERR_ARG, "ns2d needs a 2D plan with nf_max >= 4");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d_rk<float>(p, scheme, stage, S, tmp, t1, beta, visc,
                                          cfl, dt_max, s)
                      : do_ns2d_rk<double>(p, scheme, stage, S, tmp, t1, beta,
                                           visc, cfl, dt_max, s);
}

int fsx_ns2d_rk_stage_ext(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                          void* t1, double beta, const double* visc, double cfl,
                          double dt_max, const void* ext_fft, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2 || p->nf_max < 4) return fail(FSX_ERR_ARG, "ns2d needs a 2D plan with nf_max >= 4");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d_rk<float>(p, scheme, stage, S, tmp, t1, beta, visc,
                                          cfl, dt_max, s, ext_fft)
                      : do_ns2d_rk<double>(p, scheme, stage, S, tmp, t1, beta,
                                           visc, cfl, dt_max, s, ext_fft);
}

int fsx_ns3d_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                      void* t1, int has_f, double f, int project, int no_vz_kz0,
                      const double* visc, double cfl, double dt_max,
                      void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3 || p->nf_max < 6) return fail(FSX_ERR_ARG, "ns3d needs a 3D plan with nf_max >= 6");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  if (project < 0 || project > 3) return fail(FSX_ERR_ARG, "bad projection kind");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4
             ? do_ns3d_rk<float>(p, scheme, stage, S, tmp, t1, has_f, f, project,
                                 no_vz_kz0, visc, cfl, dt_max, s)
             : do_ns3d_rk<double>(p, scheme, stage, S, tmp, t1, has_f, f, project,
                                  no_vz_kz0, visc, cfl, dt_max, s);
}

int fsx_strat_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp,
                       void* t1, int has_f, double f, double N, int project,
                       int no_vz_kz0, const double* visc, double cfl,
                       double dt_max, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || !visc) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3 || p->nf_max < 7)
    return fail(FSX_ERR_ARG, "ns3d.strat needs a 3D plan with nf_max >= 7");
  if (scheme != 4 && scheme != 2) return fail(FSX_ERR_ARG, "scheme must be 2 or 4");
  if (stage < 0 || stage >= scheme) return fail(FSX_ERR_ARG, "bad stage");
  if (scheme == 4 && !tmp) return fail(FSX_ERR_ARG, "RK4 needs tmp");
  if (project < 0 || project > 3) return fail(FSX_ERR_ARG, "bad projection kind");
  hipStream_t s = (hipStream_t)stream;
  void* src = stage == 0 ? S : t1;
  return p->prec == 4
             ? do_strat<float>(p, scheme, stage, S, tmp, t1, nullptr, src, has_f, f,
                               N * N, project, no_vz_kz0, visc, cfl, dt_max, false, s)
             : do_strat<double>(p, scheme, stage, S, tmp, t1, nullptr, src, has_f,
                                f, N * N, project, no_vz_kz0, visc, cfl, dt_max,
                                false, s);
}

int fsx_strat_tendency(fsx_plan p, const void* state, void* out_fft, int has_f,
                       double f, double N, int project, int no_vz_kz0,
                       int want_maxima, void* stream) {
  PlanScope scope_(p);
  if (!p || !state || !out_fft) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 3 || p->nf_max < 7)
    return fail(FSX_ERR_ARG, "ns3d.strat needs a 3D plan with nf_max >= 7");
  if (project < 0 || project > 3) return fail(FSX_ERR_ARG, "bad projection kind");
  hipStream_t s = (hipStream_t)stream;
  const double visc[4] = {0, 0, 0, 0};
  void* st = const_cast<void*>(state);
  // stage -1: the fused pass stores the tendency instead of a stage update
  return p->prec == 4
             ? do_strat<float>(p, 4, -1, st, nullptr, st, out_fft, state, has_f, f,
                               N * N, project, no_vz_kz0, visc, 0.0, 0.0,
                               want_maxima != 0, s)
             : do_strat<double>(p, 4, -1, st, nullptr, st, out_fft, state, has_f,
                                f, N * N, project, no_vz_kz0, visc, 0.0, 0.0,
                                want_maxima != 0, s);
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

static unsigned grid_blocks(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

static KGrid kgrid(const fsx_plan_s* p) {
  KGrid g;
  g.nz = p->nz;
  g.ny = p->ny;
  g.nkx = p->nkx;
  g.dkx = p->dkx;
  g.dky = p->dky;
  g.dkz = p->dkz;
  return g;
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

namespace {

// LDS of one block record is kept within 64 KiB (two blocks per CU)
constexpr int kDiagLdsDoubles = 8192;

#ifndef FSX_DIAG_COPIES
#define FSX_DIAG_COPIES 4
#endif
#ifndef FSX_DIAG_LDS
#define FSX_DIAG_LDS (80 * 1024)
#endif

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
  // FSX_DIAG_LDS bytes (two blocks per CU at the default)
  const int span = g.off_sum - g.off_k3;
  g.ncopy = 1;
  if (span > 0)
    for (int c = FSX_DIAG_COPIES; c > 1; c >>= 1)
      if ((size_t)(g.rec + (c - 1) * span) * sizeof(double) <= FSX_DIAG_LDS) {
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
  static size_t lds_set = 0;
  if (lds > lds_set) {
    int rc = set_lds(k_diag<T, KIND>, lds);
    if (rc) return rc;
    lds_set = lds;
  }
  int rc = tm_pre(s);
  if (rc) return rc;
  hipLaunchKernelGGL((k_diag<T, KIND>), dim3(nb), dim3(256), lds, s, in, box, g, part);
  FSX_HIP(hipGetLastError());
  return tm_post(s);
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

int grow(double*& buf, size_t& cap, size_t n) {
  if (cap >= n) return 0;
  if (buf) (void)hipFree(buf);
  buf = nullptr;
  cap = 0;
  FSX_HIP(hipMalloc((void**)&buf, n * sizeof(double)));
  cap = n;
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
  const int per_q = diag_per_quantity(g);
  if (per_q > kDiagLdsDoubles) return fail(FSX_ERR_ARG, "grid too large for the LDS histograms");
  const int Q = o.Q;
  const int qstep = std::max(1, std::min(Q, kDiagLdsDoubles / per_q));
  const size_t nb = (size_t)g.nz * g.nchunk;
  DiagGeom gmax = g;
  rc = grow(p->d_dpart, p->dpart_cap, nb * (size_t)diag_record(gmax, 0, qstep));
  if (rc) return rc;
  rc = grow(p->d_dout, p->dout_cap, (size_t)total);
  if (rc) return rc;
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

int fsx_ns2d_batch_rk_stage(fsx_plan p, int scheme, int stage, void* S, void* tmp, void* t1,
                            double beta, void* stream) {
  PlanScope scope_(p);
  if (!p || !S || !t1 || (scheme == 4 && !tmp)) return fail(FSX_ERR_ARG, "bad argument");
  if (p->nb < 1 || !p->d_bpar) return fail(FSX_ERR_ARG, "not a batch plan");
  if (!((scheme == 4 && stage >= 0 && stage < 4) || (scheme == 2 && stage >= 0 && stage < 2)))
    return fail(FSX_ERR_ARG, "scheme must be 4 (stages 0-3) or 2 (stages 0-1)");
  hipStream_t s = (hipStream_t)stream;
  return p->prec == 4 ? do_ns2d_batch_rk<float>(p, scheme, stage, S, tmp, t1, beta, s)
                      : do_ns2d_batch_rk<double>(p, scheme, stage, S, tmp, t1, beta, s);
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

namespace {

bool cnn_channels_ok(int c) { return c == 8 || c == 32; }

int cnn_grid(long long n) {
  long long b = (n + 255) / 256;
  return (int)std::max<long long>(1, std::min<long long>(b, 4096));
}

// fuse: 0 -> FP32 pre-activations y; 1 -> BF16 GELU activations a (32
// channels, co_out = 32); 2 -> both
template <int CI, int FUSE, bool XF32>
void conv_fwd_launch(unsigned nblk, size_t lds, hipStream_t s, const uint16_t* x,
                     const uint16_t* w, const float* bias, float* y, int H, int W, int co_out,
                     int R, uint16_t* a) {
  hipLaunchKernelGGL((k_conv_fwd<CI, FUSE, XF32>), dim3(nblk), dim3(256), lds, s, x, w, bias, y,
                     H, W, co_out, R, a);
}

// x_f32: x is one FP32 channel (ci = 8 layer: padded while staged)
int cnn_conv_fwd(const void* x, int B, int H, int W, int ci, const void* w, const float* bias,
                 int co_out, float* y, hipStream_t s, int fuse = 0, void* a = nullptr,
                 bool x_f32 = false) {
  if (!x || !w || !bias || B < 1 || fuse < 0 || fuse > 2) return fail(FSX_ERR_ARG, "bad argument");
  if ((fuse != 1 && !y) || (fuse != 0 && !a)) return fail(FSX_ERR_ARG, "bad argument");
  if (!cnn_channels_ok(ci)) return fail(FSX_ERR_ARG, "input channels must be 8 or 32");
  if (co_out < 1 || co_out > 32) return fail(FSX_ERR_ARG, "output channels must be 1..32");
  if (fuse && co_out != 32) return fail(FSX_ERR_ARG, "fused activations need 32 output channels");
  if (x_f32 && (ci != 8 || fuse != 1))
    return fail(FSX_ERR_ARG, "an FP32 input is one channel (padded to 8) of a hidden layer");
  if (H < kCnnRows || H % kCnnRows || W < 32 || W % 32)
    return fail(FSX_ERR_ARG, "H must be a multiple of 4 and W a multiple of 32");
  // rows per workgroup: 4 when there are enough workgroups, fewer for
  // small batches (more workgroups, smaller patches)
  int R = kCnnRows;
  while (R > 1 && (long long)B * (H / R) < 512) R >>= 1;
  const size_t lds = conv_fwd_lds(ci, W, R);
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "image too wide for the convolution tile");
  const unsigned nblk = (unsigned)(B * (H / R));
  const auto* xi = reinterpret_cast<const uint16_t*>(x);
  const auto* wi = reinterpret_cast<const uint16_t*>(w);
  auto* ai = reinterpret_cast<uint16_t*>(a);
  int rc;
#define FSX_CONV_FWD(CI_, F_, XF_)                                                      \
  if (ci == CI_ && fuse == F_ && x_f32 == XF_) {                                          \
    if ((rc = set_lds(k_conv_fwd<CI_, F_, XF_>, conv_fwd_lds(CI_, W, kCnnRows)))) return rc; \
    if ((rc = tm_pre(s))) return rc;                                                      \
    conv_fwd_launch<CI_, F_, XF_>(nblk, lds, s, xi, wi, bias, y, H, W, co_out, R, ai);    \
  }
  FSX_CONV_FWD(8, 0, false)
  FSX_CONV_FWD(8, 1, false)
  FSX_CONV_FWD(8, 2, false)
  FSX_CONV_FWD(8, 1, true)
  FSX_CONV_FWD(32, 0, false)
  FSX_CONV_FWD(32, 1, false)
  FSX_CONV_FWD(32, 2, false)
#undef FSX_CONV_FWD
  FSX_HIP(hipGetLastError());
  return tm_post(s);
}

int cnn_act(const float* y, long long npix, int cin, int cpad, int use_gelu, void* a,
            hipStream_t s) {
  if (!y || !a || npix < 1 || cin < 1 || cpad < cin) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_cnn_act, dim3(cnn_grid(npix * cpad)), dim3(256), 0, s, y, npix, cin,
                     cpad, use_gelu, reinterpret_cast<uint16_t*>(a));
  FSX_HIP(hipGetLastError());
  return 0;
}

// two BF16 activation buffers of 32 channels
long long cnn_forward_work(int B, int H, int W) { return 2LL * B * H * W * 32 * 2; }

// Network forward: hidden layers write their BF16 GELU activations from the
// convolution epilogue (ping-pong between the two buffers of work), the last
// layer writes its FP32 outputs to y.
// x_f32: x is one FP32 channel (the first layer pads it to 8 while staging)
int cnn_forward(int nl, const int* cin, const int* cout, const void* const* w,
                const float* const* b, const void* x, int B, int H, int W, float* y,
                void* work, hipStream_t s, bool x_f32 = false) {
  const long long npix = (long long)B * H * W;
  uint16_t* act[2] = {reinterpret_cast<uint16_t*>(work),
                      reinterpret_cast<uint16_t*>(work) + npix * 32};
  const void* in = x;
  for (int l = 0; l < nl; ++l) {
    const bool last = l == nl - 1;
    if (!cnn_channels_ok(cin[l])) return fail(FSX_ERR_ARG, "layer input channels must be 8 or 32");
    if (!w[l] || !b[l]) return fail(FSX_ERR_ARG, "bad argument");
    if (!last && cout[l] != 32) return fail(FSX_ERR_ARG, "hidden layers have 32 channels");
    // timing: hidden 32 -> 32 layers as their own kind, work in flops
    if (l > 0 && !last)
      note(K_CNN, 2.0 * npix * 32 * kCnnTaps * 32);
    else
      note(K_OTHER, 0.0);
    if (x_f32 && l == 0 && last) return fail(FSX_ERR_ARG, "an FP32 input needs a hidden first layer");
    const int rc = last ? cnn_conv_fwd(in, B, H, W, cin[l], w[l], b[l], cout[l], y, s)
                        : cnn_conv_fwd(in, B, H, W, cin[l], w[l], b[l], 32, nullptr, s, 1,
                                       act[l & 1], x_f32 && l == 0);
    if (rc) return rc;
    in = act[l & 1];
  }
  return 0;
}

// workspace of fsx_ns2d_closure_term: vorticity and network output (FP32),
// the network's activations
long long closure_work(const fsx_plan_s* p) {
  const long long npix = (long long)p->ny * p->nx;
  return npix * (4 + 4) + cnn_forward_work(1, p->ny, p->nx);
}

}  // namespace

extern "C" {

int fsx_cnn_probe(const void* A, const void* B, float* D, void* stream) {
  if (!A || !B || !D) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     reinterpret_cast<const uint16_t*>(A), reinterpret_cast<const uint16_t*>(B),
                     D);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_probe16(const void* A, const void* B, float* D, void* stream) {
  if (!A || !B || !D) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_mfma16_probe, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     reinterpret_cast<const uint16_t*>(A), reinterpret_cast<const uint16_t*>(B),
                     D);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_conv_fwd(const void* x, int B, int H, int W, int ci, const void* w,
                     const float* bias, int co_out, float* y, void* stream) {
  return cnn_conv_fwd(x, B, H, W, ci, w, bias, co_out, y, (hipStream_t)stream);
}

int fsx_cnn_conv_fwd_act(const void* x, int B, int H, int W, int ci, const void* w,
                         const float* bias, float* y, void* a, void* stream) {
  return cnn_conv_fwd(x, B, H, W, ci, w, bias, 32, y, (hipStream_t)stream, y ? 2 : 1, a);
}

int fsx_cnn_conv_flip(const void* w, int ci, int co, int co_pad, void* wt, void* stream) {
  if (!w || !wt || ci < 1 || ci > 32 || co < 1 || co > co_pad || !cnn_channels_ok(co_pad))
    return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_conv_flip, dim3(cnn_grid(32LL * kCnnTaps * co_pad)), dim3(256), 0,
                     (hipStream_t)stream, reinterpret_cast<const uint16_t*>(w), ci, co, co_pad,
                     reinterpret_cast<uint16_t*>(wt));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_wgrad_workspace(int B, int H, int ci, long long* bytes) {
  if (!bytes || B < 1 || H < kCnnRows || !cnn_channels_ok(ci)) return fail(FSX_ERR_ARG, "bad argument");
  const long long nblk = (long long)B * (H / kCnnRows);
  const long long groups = std::min<long long>(nblk, 256);
  // weight partials, then bias partials
  *bytes = groups * 32LL * kCnnTaps * ci * 4 + 1024LL * 32 * 4;
  return 0;
}

int fsx_cnn_conv_wgrad(const void* x, int ci, const void* dy, int co_in, int B, int H, int W,
                       float* dw, float* dbias, void* work, long long work_bytes,
                       void* stream) {
  if (!x || !dy || !dw || !dbias || !work || B < 1) return fail(FSX_ERR_ARG, "bad argument");
  if (!cnn_channels_ok(ci) || !cnn_channels_ok(co_in))
    return fail(FSX_ERR_ARG, "channels must be 8 or 32");
  if (H < kCnnRows || H % kCnnRows || W < 8 || W % 8)
    return fail(FSX_ERR_ARG, "H must be a multiple of 4 and W of 8");
  long long need = 0;
  fsx_cnn_wgrad_workspace(B, H, ci, &need);
  if (work_bytes < need) return fail(FSX_ERR_ARG, "workspace too small");
  const size_t lds = conv_wgrad_lds(ci, W);
  if (lds > kLdsMax) return fail(FSX_ERR_ARG, "image too wide for the weight-gradient tile");
  hipStream_t s = (hipStream_t)stream;
  const long long nblk = (long long)B * (H / kCnnRows);
  const int groups = (int)std::min<long long>(nblk, 256);
  float* part = reinterpret_cast<float*>(work);
  const auto* xi = reinterpret_cast<const uint16_t*>(x);
  const auto* di = reinterpret_cast<const uint16_t*>(dy);
  int rc;
#define FSX_WG(CI_, CO_)                                                              \
  if ((rc = set_lds(k_conv_wgrad<CI_, CO_>, lds))) return rc;                          \
  hipLaunchKernelGGL((k_conv_wgrad<CI_, CO_>), dim3(groups), dim3(256), lds, s, xi, di, \
                     part, B, H, W);
  if (ci == 8 && co_in == 8) {
    FSX_WG(8, 8)
  } else if (ci == 8) {
    FSX_WG(8, 32)
  } else if (co_in == 8) {
    FSX_WG(32, 8)
  } else {
    FSX_WG(32, 32)
  }
#undef FSX_WG
  FSX_HIP(hipGetLastError());
  const int n = 32 * kCnnTaps * ci;
  hipLaunchKernelGGL(k_cnn_reduce, dim3((n + 63) / 64), dim3(256), 0, s, part, groups, n, dw);
  FSX_HIP(hipGetLastError());
  float* bpart = part + (size_t)groups * n;
  const long long npix = (long long)B * H * W;
  hipLaunchKernelGGL(k_cnn_bias_grad, dim3(1024), dim3(256), 0, s, di, npix, co_in, bpart);
  FSX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cnn_reduce, dim3(1), dim3(256), 0, s, bpart, 1024, 32, dbias);
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_act(const float* y, long long npix, int cin, int cpad, int use_gelu, void* a,
                void* stream) {
  return cnn_act(y, npix, cin, cpad, use_gelu, a, (hipStream_t)stream);
}

int fsx_cnn_act_bwd(const float* g, int gs, const float* y, long long npix, int c, void* d,
                    void* stream) {
  if (!g || !y || !d || npix < 1 || c < 1 || gs < c) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_cnn_act_bwd, dim3(cnn_grid(npix * c)), dim3(256), 0,
                     (hipStream_t)stream, g, gs, y, npix, c, reinterpret_cast<uint16_t*>(d));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_mse(const float* out, const float* target, long long npix, int cpad, void* d,
                double* work, double* host_loss, void* stream) {
  if (!out || !target || !d || !work || npix < 1 || cpad < 1) return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cnn_mse, dim3(256), dim3(256), 0, s, out, target, npix, cpad,
                     reinterpret_cast<uint16_t*>(d), work);
  FSX_HIP(hipGetLastError());
  if (host_loss) {
    double h[256];
    FSX_HIP(hipMemcpyAsync(h, work, sizeof(h), hipMemcpyDeviceToHost, s));
    FSX_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (int i = 0; i < 256; ++i) t += h[i];
    *host_loss = t;
  }
  return 0;
}

int fsx_cnn_adam(float* w, const float* g, float* m1, float* m2, int n, double lr, double b1,
                 double b2, double eps, int step, void* wb, void* stream) {
  if (!w || !g || !m1 || !m2 || n < 1 || step < 1) return fail(FSX_ERR_ARG, "bad argument");
  const float c1 = (float)(1.0 - std::pow(b1, step)), c2 = (float)(1.0 - std::pow(b2, step));
  hipLaunchKernelGGL(k_cnn_adam, dim3(cnn_grid(n)), dim3(256), 0, (hipStream_t)stream, w, g,
                     m1, m2, n, (float)lr, (float)b1, (float)b2, (float)eps, c1, c2,
                     reinterpret_cast<uint16_t*>(wb));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_cnn_to_bf16(const float* in, int n, void* out, void* stream) {
  if (!in || !out || n < 1) return fail(FSX_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_cnn_to_bf16, dim3(cnn_grid(n)), dim3(256), 0, (hipStream_t)stream, in,
                     n, reinterpret_cast<uint16_t*>(out));
  FSX_HIP(hipGetLastError());
  return 0;
}

int fsx_ns2d_closure_workspace(fsx_plan p, long long* bytes) {
  if (!p || !bytes) return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2) return fail(FSX_ERR_ARG, "the closure term needs a 2D plan");
  *bytes = closure_work(p);
  return 0;
}

int fsx_ns2d_closure_term(fsx_plan p, int nl, const int* cin, const int* cout,
                          const void* const* w, const float* const* b, const void* rot_fft,
                          void* out_fft, void* work, long long work_bytes, void* stream) {
  PlanScope scope_(p);
  if (!p || nl < 1 || nl > 16 || !cin || !cout || !w || !b || !rot_fft || !out_fft || !work)
    return fail(FSX_ERR_ARG, "bad argument");
  if (p->dim != 2) return fail(FSX_ERR_ARG, "the closure term needs a 2D plan");
  if (p->prec != 4) return fail(FSX_ERR_ARG, "the learned closure runs on float32 plans");
  if (cin[0] != 8 || cout[nl - 1] != 1)
    return fail(FSX_ERR_ARG, "the network maps 1 input channel (padded to 8) to 1 output");
  if (work_bytes < closure_work(p)) return fail(FSX_ERR_ARG, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  const long long npix = (long long)p->ny * p->nx;
  char* wb = reinterpret_cast<char*>(work);
  float* X = reinterpret_cast<float*>(wb);
  float* Y = reinterpret_cast<float*>(wb + npix * 4);
  void* fw = wb + npix * 8;
  int rc = do_ifft<float>(p, rot_fft, X, 1, s);
  if (!rc) rc = cnn_forward(nl, cin, cout, w, b, X, 1, p->ny, p->nx, Y, fw, s, true);
  if (!rc) rc = do_fft<float>(p, Y, out_fft, 1, s);
  if (!rc) rc = fsx_dealias(p, out_fft, 1, 0, stream);
  return rc;
}

int fsx_cnn_forward(int nl, const int* cin, const int* cout, const void* const* w,
                    const float* const* b, const void* x, int B, int H, int W, float* y,
                    void* work, long long work_bytes, void* stream) {
  if (nl < 1 || nl > 16 || !cin || !cout || !w || !b || !x || !y || !work)
    return fail(FSX_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  if (work_bytes < cnn_forward_work(B, H, W)) return fail(FSX_ERR_ARG, "workspace too small");
  return cnn_forward(nl, cin, cout, w, b, x, B, H, W, y, work, s);
}

}  // extern "C"
