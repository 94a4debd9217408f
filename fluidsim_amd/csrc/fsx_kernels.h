// Pass kernels of the fluidsim_amd engine.
//
// Array layouts in HBM (C order, complex = interleaved re/im):
//   X  (physical)     : (nz, ny, nx) real            (2D: nz = 1)
//   K  (spectral)     : (nz, ny, nkx) complex, nkx = nx/2 + 1
//   W1 (after z-pass) : (nz, nyk, ld1) complex       (3D only)
//   W2 (after y-pass) : (nz, ny,  ld2) complex
// mx is the number of kept kx indices and nyk the number of kept ky indices
// of the pass (everything for plain transforms, the bounding box of the
// non-dealiased modes for the solver passes, whose pruned modes are zero and
// are neither loaded nor stored).  ld1, ld2 >= mx are row strides rounded up
// to 128 bytes so that every run of a tile starts on a cache line.
//
// A strided pass transforms along y or z: a workgroup owns L adjacent lines
// (consecutive kx, the contiguous index) of length N, so every global access
// is a contiguous run of L complex numbers.  The first Stockham stage reads
// its butterfly inputs straight from HBM into registers, the middle stages
// go through LDS, the last stage writes straight to HBM.  A row pass
// transforms along x (real <-> half complex through a complex FFT of length
// nx/2) and is where the physical-space products are fused.
#pragma once

#include "fsx_fft.h"
#include "fsx_pointwise.h"

namespace fsx {

template <typename T>
struct StridedDesc {
  int N, logN;    // FFT length
  int L, logL;    // lines per tile (power of two)
  int nouter;     // number of independent planes
  int nlines;     // lines per plane
  int nf;         // fields transformed one after the other
  int kio;        // tiles remapped per XCD (K rows, or runs under a cache line)
  const cpx<T>* tw;  // exp(-2 pi i m / N), m < N (copied to LDS)
  // mixed-radix stages of a length that is not a power of two (gen_fft)
  int nst;
  int rad[kMaxStages];
};

// Hook of tile_fft run after the first stage (default: nothing).
struct NoFetch {
  __device__ __forceinline__ void operator()(int, int, int) const {}
};

// Stage plan of tile_fft_p2 for a length 2^logN with EPT elements per
// thread: log2 of the first stage's radix (0: a single stage of radix
// 2^logN) and of the last stage's; the stages between are radix 8.  With
// 16 elements per thread a radix-16 first stage (and a radix-16 last one
// where four factors are left) replaces a radix-8 stage plus a radix-2 or
// -4 one wherever logN is not a multiple of 3: one LDS round trip and its
// two barriers fewer per tile (1024 = 16 8 8, 2048 = 16 8 16, 256 = 16 16).
template <int EPT>
__device__ __forceinline__ void tile_plan(int logN, int& lfirst, int& llast) {
  if (logN <= 3 || (EPT >= 16 && logN == 4)) {
    lfirst = 0;
    llast = logN;
    return;
  }
  const bool r16 = EPT >= 16 && logN % 3 != 0;
  lfirst = r16 ? 4 : 3;
  const int lmax = r16 ? 4 : 3;
  int logNs = lfirst;
  while (logN - logNs > lmax) logNs += 3;
  llast = logN - logNs;
}

// Full transform of one tile.  src(e, l, p) gives the input at position p
// of local line l (e = register slot of the first stage, stable across
// calls); dst(e, l, p, v) receives the output at position p (e = register
// slot of the last stage).  fetch(e, l, p) is called for every first-stage
// slot once that slot's input has been consumed: a multi-transform pass
// issues there the global loads of its next transform (same slots), so
// they are in flight during this transform's LDS stages.  Ends with a
// barrier so buf can be reused.
template <typename T, int DIR, int EPT, int kT, class Src, class Dst,
          class Fetch = NoFetch>
__device__ __forceinline__ void tile_fft_p2(cpx<T>* buf, const cpx<T>* tw,
                                            int logN, int logL, Src&& src,
                                            Dst&& dst, const Fetch& fetch = Fetch()) {
  const int lmask = (1 << logL) - 1;
  int lfirst, llast;
  tile_plan<EPT>(logN, lfirst, llast);
  if (lfirst == 0) {
    // single stage: HBM -> registers -> HBM
    auto single = [&](auto Rc) {
      constexpr int R = decltype(Rc)::value;
      constexpr int NB = EPT / R;
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int b = threadIdx.x + i * kT;
        if (b <= lmask) {
          cpx<T> v[R];
#pragma unroll
          for (int r = 0; r < R; ++r) v[r] = src(i * R + r, b, r);
          dftR<T, DIR, R>(v);
#pragma unroll
          for (int r = 0; r < R; ++r) dst(i * R + r, b, r, v[r]);
#pragma unroll
          for (int r = 0; r < R; ++r) fetch(i * R + r, b, r);
        }
      }
    };
    if constexpr (EPT >= 16) {
      if (logN == 4) single(std::integral_constant<int, 16>());
    }
    if (logN == 3)
      single(std::integral_constant<int, 8>());
    else if (logN == 2)
      single(std::integral_constant<int, 4>());
    else if (logN < 2)
      single(std::integral_constant<int, 2>());
    __syncthreads();
    return;
  }
  // first stage, Ns = 1: inputs from HBM, outputs to LDS
  auto first = [&](auto Rc) {
    constexpr int R = decltype(Rc)::value;
    constexpr int LOGR = R == 16 ? 4 : 3;
    constexpr int NB = EPT / R;
    const int lognb = logN - LOGR;
    const int total = 1 << (lognb + logL);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = threadIdx.x + i * kT;
      if (b < total) {
        const int line = b & lmask;
        const int j = b >> logL;
        cpx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src(i * R + r, line, j + (r << lognb));
        dftR<T, DIR, R>(v);
#pragma unroll
        for (int r = 0; r < R; ++r) buf[lpad((((j << LOGR) + r) << logL) + line)] = v[r];
#pragma unroll
        for (int r = 0; r < R; ++r) fetch(i * R + r, line, j + (r << lognb));
      }
    }
    __syncthreads();
  };
  if constexpr (EPT >= 16) {
    if (lfirst == 4) first(std::integral_constant<int, 16>());
  }
  if (lfirst == 3) first(std::integral_constant<int, 8>());
  int logNs = lfirst;
  while (logN - logNs > llast) {
    lds_stage_lf<T, DIR, EPT, 8, kT>(buf, logN, logNs, logL, tw);
    logNs += 3;
  }
  // last stage: k = j, outputs at j + r Ns, to HBM
  auto last = [&](auto Rc) {
    constexpr int R = decltype(Rc)::value;
    constexpr int LOGR = R == 16 ? 4 : (R == 8 ? 3 : (R == 4 ? 2 : 1));
    constexpr int NB = EPT / R;
    const int lognb = logN - LOGR;
    const int total = 1 << (lognb + logL);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = threadIdx.x + i * kT;
      if (b < total) {
        const int line = b & lmask;
        const int j = b >> logL;
        cpx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = buf[lpad(((j + (r << lognb)) << logL) + line)];
        twiddle<T, DIR, R, LOGR>(v, j, logN, logN - LOGR, tw, 0);
        dftR<T, DIR, R>(v);
#pragma unroll
        for (int r = 0; r < R; ++r) dst(i * R + r, line, j + (r << lognb), v[r]);
      }
    }
  };
  if constexpr (EPT >= 16) {
    if (llast == 4) last(std::integral_constant<int, 16>());
  }
  if (llast == 3)
    last(std::integral_constant<int, 8>());
  else if (llast == 2)
    last(std::integral_constant<int, 4>());
  else if (llast == 1)
    last(std::integral_constant<int, 2>());
  __syncthreads();
}

// Full transform of one tile of 2^logL lines of length d.N.  GEN = false:
// power-of-two length (tile_fft_p2, registers at both ends).  GEN = true:
// any length (gen_fft); slot e of both src and dst is the tile element
// threadIdx.x + e kT (line-fast), and the tile uses two LDS buffers.
template <typename T, int DIR, int EPT, int kT, bool GEN, class Src, class Dst,
          class Fetch = NoFetch>
__device__ __forceinline__ void tile_fft(cpx<T>* buf, const cpx<T>* tw,
                                         const StridedDesc<T>& d, int logL, Src&& src,
                                         Dst&& dst, const Fetch& fetch = Fetch()) {
  if constexpr (!GEN) {
    tile_fft_p2<T, DIR, EPT, kT>(buf, tw, d.logN, logL, src, dst, fetch);
  } else {
    const int tot = d.N << logL;
    const int lmask = (1 << logL) - 1;
    cpx<T>* A = buf;
    cpx<T>* B = buf + tot + (tot >> 5) + 1;
    const TileLayout lay{logL, 0};
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = threadIdx.x + e * kT;
      if (idx < tot) A[lpad(idx)] = src(e, idx & lmask, idx >> logL);
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = threadIdx.x + e * kT;
      if (idx < tot) fetch(e, idx & lmask, idx >> logL);
    }
    __syncthreads();
    const cpx<T>* res = gen_fft<T, DIR, kT>(A, B, lay, tw, d.N, 1 << logL, d.nst, d.rad);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = threadIdx.x + e * kT;
      if (idx < tot) dst(e, idx & lmask, idx >> logL, res[lpad(idx)]);
    }
    __syncthreads();
  }
}

// Position of register slot e of tile_fft's last stage: returns false when
// the slot holds no butterfly output, else its local line and position.
template <int EPT, int kT, bool GEN, typename T>
__device__ __forceinline__ bool last_slot(int e, const StridedDesc<T>& d, int logL,
                                          int& line, int& p) {
  if constexpr (GEN) {
    const int idx = threadIdx.x + e * kT;
    if (idx >= (d.N << logL)) return false;
    line = idx & ((1 << logL) - 1);
    p = idx >> logL;
    return true;
  }
  const int logN = d.logN;
  int lfirst, logR;
  tile_plan<EPT>(logN, lfirst, logR);
  const int i = e >> logR, r = e & ((1 << logR) - 1);
  if (i >= (EPT >> logR)) return false;
  const int b = threadIdx.x + i * kT;
  const int lognb = logN - logR;
  if (b >= (1 << (lognb + logL))) return false;
  line = b & ((1 << logL) - 1);
  p = (b >> logL) + (r << lognb);
  return true;
}

// Elements per thread of the multi-transform z-passes (tile = kThr x EPT /
// div): the curl pass stashes 3 and the Runge-Kutta passes 3-4 transforms
// in registers, so their tiles are halved by default.
constexpr int kZCurlEptDiv = 2;
constexpr int kZRKEptDiv = 2;

// ---------------------------------------------------------------- strided
// Workgroups are dispatched round-robin over the XCDs (workgroup b runs on
// XCD b mod kNXCD), each XCD with its own L2.  In passes over K rows
// (nkx = nx/2 + 1 complex values, not line-aligned), and in passes whose
// tiles hold runs shorter than a 128-byte line (long transforms leave few
// lines per tile), neighbouring line chunks share cache lines, so there
// consecutive tiles are given to the same XCD: XCD x takes the x-th
// contiguous range.
constexpr int kNXCD = 8;
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  if (nb < kNXCD) return b;
  const int x = b % kNXCD, local = b / kNXCD;
  const int q = nb / kNXCD, rem = nb % kNXCD;
  return x < rem ? x * (q + 1) + local : rem * (q + 1) + (x - rem) * q + local;
}

template <typename T, int DIR, int EPT, int kT, class Op, bool GEN = false>
static __global__ __launch_bounds__(kT) void k_strided(Op op) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const StridedDesc<T>& d = op.d;
  cpx<T>* s_tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* buf = s_tw + d.N;
  for (int i = threadIdx.x; i < d.N; i += kT) s_tw[i] = d.tw[i];
  // (tile_fft's first barrier precedes every twiddle read)
  const int nchunks = (d.nlines + d.L - 1) >> d.logL;
  const int b = d.kio ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int o = b / nchunks;
  const int line0 = (b - o * nchunks) << d.logL;
  op.template run<DIR, EPT, kT, GEN>(buf, s_tw, o, line0);
}

// Plain pass: one field per workgroup (blockIdx.y), load()/store().
template <typename T, class Self>
struct PlainRun {
  static constexpr int kEptDiv = 1;
  static constexpr bool kFieldPerBlock = true;
  template <int DIR, int EPT, int kT, bool GEN>
  __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int o, int line0) const {
    const Self& s = static_cast<const Self&>(*this);
    const int nlines = s.d.nlines;
    const int f = blockIdx.y;
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, s.d, s.d.logL,
        [&](int, int l, int p) {
          const int line = line0 + l;
          return line < nlines ? s.load(f, o, line, p) : zero<T>();
        },
        [&](int, int l, int p, cpx<T> v) {
          const int line = line0 + l;
          if (line < nlines) s.store(f, o, line, p, v);
        });
  }
};

// Plain z-pass, inverse: K (nz, ny, nkx) -> W1 (nz, nyk, ld1).
// Lines are kx in [0, mx); planes are the kept ky rows.
template <typename T>
struct OpZInv : PlainRun<T, OpZInv<T>> {
  StridedDesc<T> d;
  const cpx<T>* in;
  cpx<T>* out;
  long long ifs, ofs;  // field strides (elements)
  int ny, nkx, ld1;
  AxisMap ym, zm;
  __device__ cpx<T> load(int f, int o, int l, int p) const {
    if (zm.compact(p) < 0) return zero<T>();
    return in[f * ifs + (p * ny + ym.full(o)) * nkx + l];
  }
  __device__ void store(int f, int o, int l, int p, cpx<T> v) const {
    out[f * ofs + (p * ym.nk() + o) * ld1 + l] = v;
  }
};

// y-pass, inverse: W1 (nz, nyk, ld1) -> W2 (nz, ny, ld2); planes are z.
template <typename T>
struct OpYInv3 : PlainRun<T, OpYInv3<T>> {
  StridedDesc<T> d;
  const cpx<T>* in;
  cpx<T>* out;
  long long ifs, ofs;
  int ny, ld1, ld2;
  AxisMap ym;
  __device__ cpx<T> load(int f, int z, int l, int p) const {
    const int jy = ym.compact(p);
    if (jy < 0) return zero<T>();
    return in[f * ifs + (z * ym.nk() + jy) * ld1 + l];
  }
  __device__ void store(int f, int z, int l, int p, cpx<T> v) const {
    out[f * ofs + (z * ny + p) * ld2 + l] = v;
  }
};

// y-pass, forward: W2 (nz, ny, ld2) -> W1 (nz, nyk, ld1).
template <typename T>
struct OpYFwd3 : PlainRun<T, OpYFwd3<T>> {
  StridedDesc<T> d;
  const cpx<T>* in;
  cpx<T>* out;
  long long ifs, ofs;
  int ny, ld1, ld2;
  AxisMap ym;
  __device__ cpx<T> load(int f, int z, int l, int p) const {
    return in[f * ifs + (z * ny + p) * ld2 + l];
  }
  __device__ void store(int f, int z, int l, int p, cpx<T> v) const {
    const int jy = ym.compact(p);
    if (jy >= 0) out[f * ofs + (z * ym.nk() + jy) * ld1 + l] = v;
  }
};

// z-pass, forward: W1 (nz, nyk, ld1) -> K (nz, ny, nkx); only kept modes
// are written, dealiased ones (mask) as zero.
template <typename T>
struct OpZFwd : PlainRun<T, OpZFwd<T>> {
  StridedDesc<T> d;
  const cpx<T>* in;
  cpx<T>* out;
  long long ifs, ofs;
  int ny, nkx, ld1;
  AxisMap ym, zm;
  const unsigned char* mask;
  __device__ cpx<T> load(int f, int jy, int l, int p) const {
    return in[f * ifs + (p * ym.nk() + jy) * ld1 + l];
  }
  __device__ void store(int f, int jy, int l, int p, cpx<T> v) const {
    if (zm.compact(p) < 0) return;
    const int idx = (p * ny + ym.full(jy)) * nkx + l;
    if (mask && mask[idx]) v = zero<T>();
    out[f * ofs + idx] = v;
  }
};

// 2D y-pass, inverse: K (ny, nkx) -> W2 (ny, ld2).
template <typename T>
struct OpYInv2 : PlainRun<T, OpYInv2<T>> {
  StridedDesc<T> d;
  const cpx<T>* in;
  cpx<T>* out;
  long long ifs, ofs;
  int nkx, ld2;
  AxisMap ym;
  __device__ cpx<T> load(int f, int, int l, int p) const {
    if (ym.compact(p) < 0) return zero<T>();
    return in[f * ifs + p * nkx + l];
  }
  __device__ void store(int f, int, int l, int p, cpx<T> v) const {
    out[f * ofs + p * ld2 + l] = v;
  }
};

// 2D y-pass, forward: W2 (ny, ld2) -> K (ny, nkx), kept modes only.
template <typename T>
struct OpYFwd2 : PlainRun<T, OpYFwd2<T>> {
  StridedDesc<T> d;
  const cpx<T>* in;
  cpx<T>* out;
  long long ifs, ofs;
  int nkx, ld2;
  AxisMap ym;
  const unsigned char* mask;
  __device__ cpx<T> load(int f, int, int l, int p) const {
    return in[f * ifs + p * ld2 + l];
  }
  __device__ void store(int f, int, int l, int p, cpx<T> v) const {
    if (ym.compact(p) < 0) return;
    const int idx = p * nkx + l;
    if (mask && mask[idx]) v = zero<T>();
    out[f * ofs + idx] = v;
  }
};

// ns3d z-pass, inverse, with the curl folded in.  omega = i k x v needs
// i kz v only through d/dz, which is a transform of its own along z, while
// kx and ky are constant on a z column and multiply after the transform:
//   omega_x = i ky Vz - Dz[vy],  omega_y = Dz[vx] - i kx Vz,
//   omega_z = i kx Vy - i ky Vx (+ f on the kx = ky = 0 column),
// with V = iFFT_z(v) and Dz[a] = iFFT_z(i kz a).  Five transforms per column,
// each input read from HBM once; outputs 0-2 are v, 3-5 omega.
template <typename T>
struct OpZInvCurl {
  static constexpr int kEptDiv = kZCurlEptDiv;  // register stashes: half-size tiles
  static constexpr bool kFieldPerBlock = false;
  StridedDesc<T> d;
  const cpx<T>* in;  // 3 fields
  cpx<T>* out;       // 6 fields
  long long ifs, ofs;
  int ny, nz, nkx, ld1;
  AxisMap ym, zm;
  T dkx, dky, dkz;
  int has_f;
  T fcor;
  int with_b;  // ns3d.strat: input 3 (b) transformed into output 6

  template <int DIR, int EPT, int kT, bool GEN>
  __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int o, int line0) const {
    const int nlines = d.nlines;
    const int iy = ym.full(o);
    const T ky = T(ksigned(iy, ny)) * dky;
    const int nyk = ym.nk();
    cpx<T> sVz[EPT], sv[EPT], sVy[EPT];
    auto ld = [&](int c, int l, int p) {
      const int line = line0 + l;
      if (line >= nlines || zm.compact(p) < 0) return zero<T>();
      return in[c * ifs + (p * ny + iy) * nkx + line];
    };
    auto st = [&](int c, int l, int p, cpx<T> v) {
      const int line = line0 + l;
      if (line < nlines) out[c * ofs + (p * nyk + o) * ld1 + line] = v;
    };
    auto kz_of = [&](int p) { return T(ksigned(p, nz)) * dkz; };
    // Each input is loaded once, into sv (vy, then vx) or sVz (b, once Vz
    // is no longer needed), while the transform before the one that first
    // uses it runs (tile_fft's fetch hook): only the first load of a tile
    // waits for HBM.
    // Vz; vy fetched
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int, int l, int p) { return ld(2, l, p); },
        [&](int e, int l, int p, cpx<T> v) {
          sVz[e] = v;
          st(2, l, p, v);
        },
        [&](int e, int l, int p) { sv[e] = ld(1, l, p); });
    // omega_x = i ky Vz - Dz[vy]
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int e, int, int p) { return mul_i(scale(sv[e], kz_of(p))); },
        [&](int e, int l, int p, cpx<T> v) {
          st(3, l, p, mul_i(scale(sVz[e], ky)) - v);
        });
    // Vy; vx fetched over vy once the first stage has read it
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int e, int, int) { return sv[e]; },
        [&](int e, int l, int p, cpx<T> v) {
          sVy[e] = v;
          st(1, l, p, v);
        },
        [&](int e, int l, int p) { sv[e] = ld(0, l, p); });
    // omega_y = Dz[vx] - i kx Vz
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int e, int, int p) { return mul_i(scale(sv[e], kz_of(p))); },
        [&](int e, int l, int p, cpx<T> v) {
          const T kx = T(line0 + l) * dkx;
          st(4, l, p, v - mul_i(scale(sVz[e], kx)));
        });
    // Vx and omega_z = i kx Vy - i ky Vx (+ f); b fetched into sVz
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int e, int, int) { return sv[e]; },
        [&](int e, int l, int p, cpx<T> v) {
          const T kx = T(line0 + l) * dkx;
          st(0, l, p, v);
          cpx<T> w = mul_i(scale(sVy[e], kx) - scale(v, ky));
          if (has_f && iy == 0 && line0 + l == 0) w.x += fcor;
          st(5, l, p, w);
        },
        [&](int e, int l, int p) {
          if (with_b) sVz[e] = ld(3, l, p);
        });
    if (with_b)
      tile_fft<T, DIR, EPT, kT, GEN>(
          buf, tw, d, d.logL, [&](int e, int, int) { return sVz[e]; },
          [&](int, int l, int p, cpx<T> v) { st(6, l, p, v); });
  }
};

// ns2d y-pass, inverse, from the vorticity: one read of rot, four
// transforms: ux = i ky rot / k^2, uy = -i kx rot / k^2, d_x rot = i kx rot,
// d_y rot = i ky rot (k^2 = 1e-14 at k = 0, the K2_not0 convention).
template <typename T>
struct OpYInv2NS {
  static constexpr int kEptDiv = 1;
  static constexpr bool kFieldPerBlock = false;
  // the four outputs can share one tile as 4 L lines (small transforms)
  static constexpr int kLineMul = 4;
  StridedDesc<T> d;
  const cpx<T>* in;
  cpx<T>* out;
  long long ofs;
  // plane o of a batch of independent runs: in + o * ips, out + o * ops
  long long ips = 0, ops = 0;
  int ny, nkx, ld2;
  AxisMap ym;
  T dkx, dky;

  template <int DIR, int EPT, int kT, bool GEN>
  __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int o, int line0) const {
    const int nlines = d.nlines;
    const cpx<T>* in = this->in + o * ips;
    cpx<T>* out = this->out + o * ops;
    cpx<T> sr[EPT];
    auto st = [&](int c, int l, int p, cpx<T> v) {
      const int line = line0 + l;
      if (line < nlines) out[c * ofs + p * ld2 + line] = v;
    };
    auto inv_k2 = [&](int l, int p) {
      const T kx = T(line0 + l) * dkx, ky = T(ksigned(p, ny)) * dky;
      T k2 = kx * kx + ky * ky;
      if (k2 == T(0)) k2 = T(1e-14);
      return T(1) / k2;
    };
    if (((size_t)d.N << (d.logL + 2)) <= (size_t)EPT * kT) {
      // one transform of 4 L lines: line l is output l >> logL of line
      // l & (L - 1), each line computed as in the sequential form below
      const int lm = (1 << d.logL) - 1;
      tile_fft<T, DIR, EPT, kT, GEN>(
          buf, tw, d, d.logL + 2,
          [&](int, int l, int p) {
            const int c = l >> d.logL, li = l & lm;
            const int line = line0 + li;
            cpx<T> r = zero<T>();
            if (line < nlines && ym.compact(p) >= 0) r = in[p * nkx + line];
            T f;
            if (c == 0)
              f = T(ksigned(p, ny)) * dky * inv_k2(li, p);
            else if (c == 1)
              f = -(T(line0 + li) * dkx) * inv_k2(li, p);
            else if (c == 2)
              f = T(line0 + li) * dkx;
            else
              f = T(ksigned(p, ny)) * dky;
            return mul_i(scale(r, f));
          },
          [&](int, int l, int p, cpx<T> v) { st(l >> d.logL, l & lm, p, v); });
      return;
    }
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int e, int l, int p) {
          const int line = line0 + l;
          cpx<T> r = zero<T>();
          if (line < nlines && ym.compact(p) >= 0) r = in[p * nkx + line];
          sr[e] = r;
          const T ky = T(ksigned(p, ny)) * dky;
          return mul_i(scale(r, ky * inv_k2(l, p)));
        },
        [&](int, int l, int p, cpx<T> v) { st(0, l, p, v); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int e, int l, int p) {
          const T kx = T(line0 + l) * dkx;
          return mul_i(scale(sr[e], -kx * inv_k2(l, p)));
        },
        [&](int, int l, int p, cpx<T> v) { st(1, l, p, v); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int e, int l, int) { return mul_i(scale(sr[e], T(line0 + l) * dkx)); },
        [&](int, int l, int p, cpx<T> v) { st(2, l, p, v); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int e, int, int p) { return mul_i(scale(sr[e], T(ksigned(p, ny)) * dky)); },
        [&](int, int l, int p, cpx<T> v) { st(3, l, p, v); });
  }
};

// OpYInv2NS with one output field per workgroup (blockIdx.y): four times
// the workgroups, a quarter of the serial work each, the input read once per
// output (long transforms with few lines, such as ns2d at 2048^2).
template <typename T>
struct OpYInv2NSOne : OpYInv2NS<T> {
  static constexpr bool kFieldPerBlock = true;
  static constexpr int kLineMul = 1;

  template <int DIR, int EPT, int kT, bool GEN>
  __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int o, int line0) const {
    const StridedDesc<T>& d = this->d;
    const int nlines = d.nlines;
    const int c = blockIdx.y;
    const cpx<T>* in = this->in + o * this->ips;
    cpx<T>* out = this->out + o * this->ops + c * this->ofs;
    const int ny = this->ny, nkx = this->nkx, ld2 = this->ld2;
    const T dkx = this->dkx, dky = this->dky;
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int, int l, int p) {
          const int line = line0 + l;
          cpx<T> r = zero<T>();
          if (line < nlines && this->ym.compact(p) >= 0) r = in[p * nkx + line];
          const T kx = T(line) * dkx, ky = T(ksigned(p, ny)) * dky;
          T k2 = kx * kx + ky * ky;
          if (k2 == T(0)) k2 = T(1e-14);
          const T ik2 = T(1) / k2;
          // the multipliers of OpYInv2NS's outputs 0..3
          const T f = c == 0 ? ky * ik2 : (c == 1 ? -kx * ik2 : (c == 2 ? kx : ky));
          return mul_i(scale(r, f));
        },
        [&](int, int l, int p, cpx<T> v) {
          const int line = line0 + l;
          if (line < nlines) out[p * ld2 + line] = v;
        });
  }
};

// ns3d z-pass, forward, with the Runge-Kutta stage folded in: the three
// tendency components are transformed one after the other (the first two
// kept in registers), then each kept mode is projected, masked and advanced
// (rk_point); the tendency never goes to HBM.
template <typename T>
struct OpZFwdRK3 {
  static constexpr int kEptDiv = kZRKEptDiv;
  static constexpr bool kFieldPerBlock = false;
  StridedDesc<T> d;
  const cpx<T>* in;  // W1, 3 fields
  long long ifs;
  int ny, nkx, ld1;
  AxisMap ym, zm;
  const unsigned char* mask;
  RKArgs<T> rk;

  template <int DIR, int EPT, int kT, bool GEN>
  __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int o, int line0) const {
    const int nlines = d.nlines;
    const int nyk = ym.nk();
    const int iy = ym.full(o);
    const int iky = ksigned(iy, ny);
    const double dt = *rk.dt;
    cpx<T> s0[EPT], s1[EPT], s2[EPT];
    auto ld = [&](int c, int l, int p) {
      const int line = line0 + l;
      if (line >= nlines) return zero<T>();
      return in[c * ifs + (p * nyk + o) * ld1 + line];
    };
    // input c + 1 is fetched into s(c+1) while transform c runs; the
    // stash's first stage reads it before its last stage stores the output
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int, int l, int p) { return ld(0, l, p); },
        [&](int e, int, int, cpx<T> v) { s0[e] = v; },
        [&](int e, int l, int p) { s1[e] = ld(1, l, p); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int e, int, int) { return s1[e]; },
        [&](int e, int, int, cpx<T> v) { s1[e] = v; },
        [&](int e, int l, int p) { s2[e] = ld(2, l, p); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int e, int, int) { return s2[e]; },
        [&](int e, int, int, cpx<T> v) { s2[e] = v; });
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      int l, p;
      if (!last_slot<EPT, kT, GEN>(e, d, d.logL, l, p)) continue;
      const int line = line0 + l;
      if (line >= nlines || zm.compact(p) < 0) continue;
      const int idx = (p * ny + iy) * nkx + line;
      cpx<T> N[3] = {s0[e], s1[e], s2[e]};
      if (mask && mask[idx]) N[0] = N[1] = N[2] = zero<T>();
      rk_point<T, 3>(rk, rk.visc, dt, idx, line, iky, ksigned(p, d.N), N);
    }
  }
};

// ns3d.strat z-pass, forward, with the Runge-Kutta stage folded in
// (fluidsim/solvers/ns3d/strat/solver.py:29-33, 138-216): W1 fields 0-2 are
// the transforms of v x omega, 3-5 those of v b.  Per kept mode
//   N_v = (F_x, F_y, F_z + b),  N_b = -i k.(v b)^ - N^2 v_z,
// with b and v_z of the state the tendency is evaluated on (rk.src); the
// projection acts on N_v only (rk_point).
template <typename T>
struct OpZFwdRKStrat {
  static constexpr int kEptDiv = kZRKEptDiv;
  static constexpr bool kFieldPerBlock = false;
  StridedDesc<T> d;
  const cpx<T>* in;  // W1, 6 fields
  long long ifs;
  int ny, nkx, ld1;
  AxisMap ym, zm;
  const unsigned char* mask;
  T dkx, dky, dkz;
  T N2;
  RKArgs<T> rk;

  template <int DIR, int EPT, int kT, bool GEN>
  __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int o, int line0) const {
    const int nlines = d.nlines;
    const int nyk = ym.nk();
    const int iy = ym.full(o);
    const int iky = ksigned(iy, ny);
    const T ky = T(iky) * dky;
    const double dt = *rk.dt;
    cpx<T> s0[EPT], s1[EPT], s2[EPT], sb[EPT];
    auto ld = [&](int c, int l, int p) {
      const int line = line0 + l;
      if (line >= nlines) return zero<T>();
      return in[c * ifs + (p * nyk + o) * ld1 + line];
    };
    // inputs 1-3 are fetched into s1, s2, sb while the transform before
    // runs (see OpZFwdRK3)
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int, int l, int p) { return ld(0, l, p); },
        [&](int e, int, int, cpx<T> v) { s0[e] = v; },
        [&](int e, int l, int p) { s1[e] = ld(1, l, p); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int e, int, int) { return s1[e]; },
        [&](int e, int, int, cpx<T> v) { s1[e] = v; },
        [&](int e, int l, int p) { s2[e] = ld(2, l, p); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int e, int, int) { return s2[e]; },
        [&](int e, int, int, cpx<T> v) { s2[e] = v; },
        [&](int e, int l, int p) { sb[e] = ld(3, l, p); });
    // sb = kx G_x + ky G_y + kz G_z  (N_b = -i sb - N^2 v_z)
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int e, int, int) { return sb[e]; },
        [&](int e, int l, int, cpx<T> v) { sb[e] = scale(v, T(line0 + l) * dkx); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int, int l, int p) { return ld(4, l, p); },
        [&](int e, int, int, cpx<T> v) { sb[e] = sb[e] + scale(v, ky); });
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL, [&](int, int l, int p) { return ld(5, l, p); },
        [&](int e, int, int p, cpx<T> v) {
          sb[e] = sb[e] + scale(v, T(ksigned(p, d.N)) * dkz);
        });
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      int l, p;
      if (!last_slot<EPT, kT, GEN>(e, d, d.logL, l, p)) continue;
      const int line = line0 + l;
      if (line >= nlines || zm.compact(p) < 0) continue;
      const int idx = (p * ny + iy) * nkx + line;
      const cpx<T> bs = rk.src[3][idx], vzs = rk.src[2][idx];
      cpx<T> N[4] = {s0[e], s1[e], s2[e] + bs,
                     scale(mul_i(sb[e]), T(-1)) - scale(vzs, N2)};
      if (mask && mask[idx]) N[0] = N[1] = N[2] = N[3] = zero<T>();
      rk_point<T, 4>(rk, rk.visc, dt, idx, line, iky, ksigned(p, d.N), N);
    }
  }
};

// ns2d y-pass, forward, with the Runge-Kutta stage folded in.
template <typename T>
struct OpYFwd2RK {
  static constexpr int kEptDiv = 1;
  static constexpr bool kFieldPerBlock = false;
  StridedDesc<T> d;
  const cpx<T>* in;  // W2, 1 field
  int ny, nkx, ld2;
  AxisMap ym;
  const unsigned char* mask;
  RKArgs<T> rk;
  // plane o of a batch of independent runs: W2 rows at o * ips, state modes
  // at o * ops, and the run's dt and viscosities at bpar[5 o .. 5 o + 4]
  // (dt, nu_2, nu_4, nu_8, nu_m4); bpar null: rk.dt and rk.visc
  long long ips = 0, ops = 0;
  const double* bpar = nullptr;

  template <int DIR, int EPT, int kT, bool GEN>
  __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int o, int line0) const {
    const int nlines = d.nlines;
    const cpx<T>* in = this->in + o * ips;
    double dt;
    Visc visc;
    if (bpar) {
      const double* q = bpar + 5 * o;
      dt = q[0];
      visc = Visc{q[1], q[2], q[3], q[4]};
    } else {
      dt = *rk.dt;
      visc = rk.visc;
    }
    const size_t base = (size_t)o * ops;
    tile_fft<T, DIR, EPT, kT, GEN>(
        buf, tw, d, d.logL,
        [&](int, int l, int p) {
          const int line = line0 + l;
          return line < nlines ? in[p * ld2 + line] : zero<T>();
        },
        [&](int, int l, int p, cpx<T> v) {
          const int line = line0 + l;
          if (line >= nlines || ym.compact(p) < 0) return;
          const int idx = p * nkx + line;
          cpx<T> N[1] = {v};
          if (mask && mask[idx]) N[0] = zero<T>();
          rk_point<T, 1>(rk, visc, dt, base + idx, line, ksigned(p, ny), 0, N);
        });
  }
};

// ---------------------------------------------------------------- rows
template <typename T>
struct XDesc {
  int nx, lognx, M, logM;  // M = nx / 2 complex points per row
  int RB, logRB;           // rows per workgroup (power of two)
  int nrows;
  int nf_in, nf_out;
  int mx_in, ld_in;    // kept kx and row stride of the c2r input rows
  int mx_out, ld_out;  // kept kx and row stride of the r2c output rows
  long long ifs, ofs;  // field strides (elements)
  T scale;             // applied to the r2c output
  const cpx<T>* tw;    // exp(-2 pi i m / nx), m < nx (copied to LDS)
  // mixed-radix stages of a length-nx complex transform (k_xpass_gen)
  int nst;
  int rad[kMaxStages];
};

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide max of up to 3 values, then one atomicMax per value on the bit
// pattern of a non-negative double (monotonic as unsigned 64-bit).
template <typename T, int kT>
__device__ void block_max_atomic(T m0, T m1, T m2, int nv,
                                 unsigned long long* dst) {
  __shared__ T red[3][kT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  m0 = wave_max(m0);
  m1 = wave_max(m1);
  m2 = wave_max(m2);
  if (lane == 0) {
    red[0][wid] = m0;
    red[1][wid] = m1;
    red[2][wid] = m2;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    T m = red[threadIdx.x][0];
    for (int w = 1; w < kT / 64; ++w) m = fmax(m, red[threadIdx.x][w]);
    const double md = (double)m;
    atomicMax(dst + threadIdx.x, (unsigned long long)__double_as_longlong(md));
  }
  __syncthreads();
}

// Real view of the padded row layout: real x of line lr.
__device__ __forceinline__ int rpad(int lr, int x, int ldr) {
  return 2 * (lr * ldr + lpad(x >> 1)) + (x & 1);
}

// At least 4 waves per SIMD (caps the VGPR count).
template <typename T, int EPT, int kT, class Op>
static __global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(4)))
void k_xpass(Op op) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const XDesc<T>& d = op.d;
  cpx<T>* s_tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* buf = s_tw + d.nx;
  T* bufr = reinterpret_cast<T*>(buf);
  const int M = d.M, logM = d.logM, nx = d.nx;
  const int H = M >> 1, logH = logM - 1;
  const int ldr = lpad(M);  // padded complex row stride in LDS
  const int row0 = blockIdx.x << d.logRB;
  const int tid = threadIdx.x;
  for (int i = tid; i < nx; i += kT) s_tw[i] = d.tw[i];
  __syncthreads();

  if constexpr (Op::kInv) {
    // c2r pre-processing, one thread per pair (k, M - k):
    // Z[k] = (X[k] + conj X[M-k]) + i conj(W^k) (X[k] - conj X[M-k]),
    // W = exp(-2 pi i / nx); only the real parts of X[0] and X[M] are used;
    // the k = 0 thread also does the self-paired k = M/2.  The loads of
    // kU consecutive pairs of a thread are issued before any is used, so
    // their latencies overlap.
    constexpr int kU = 4;
    const int n_in = (d.nf_in << d.logRB) << logH;
    for (int base = tid; base < n_in; base += kU * kT) {
      cpx<T> Xk[kU], Xm[kU], Xh[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int idx = base + u * kT;
        Xk[u] = zero<T>();
        Xm[u] = zero<T>();
        Xh[u] = zero<T>();
        const int k = idx & (H - 1);
        const int lr = idx >> logH;
        const int row = row0 + (lr & (d.RB - 1));
        if (idx < n_in && row < d.nrows) {
          const cpx<T>* src = op.in + (lr >> d.logRB) * d.ifs + (long long)row * d.ld_in;
          if (k < d.mx_in) Xk[u] = src[k];
          if (M - k < d.mx_in) Xm[u] = src[M - k];
          if (k == 0 && H < d.mx_in) Xh[u] = src[H];
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int idx = base + u * kT;
        if (idx >= n_in) break;
        const int k = idx & (H - 1);
        const int lr = idx >> logH;
        cpx<T>* Z = buf + lr * ldr;
        if (k == 0) {
          Xk[u].y = T(0);
          Xm[u].y = T(0);
        }
        const cpx<T> w = s_tw[k];
        const cpx<T> dk = Xk[u] - conj(Xm[u]), dm = Xm[u] - conj(Xk[u]);
        Z[lpad(k)] = (Xk[u] + conj(Xm[u])) + mul_i(conj(w) * dk);
        if (k == 0) {
          const cpx<T> dh = Xh[u] - conj(Xh[u]);
          Z[lpad(H)] = (Xh[u] + conj(Xh[u])) + mul_i(conj(s_tw[H]) * dh);
        } else {
          // conj(W^(M-k)) = -W^k
          Z[lpad(M - k)] = (Xm[u] + conj(Xk[u])) - mul_i(w * dm);
        }
      }
    }
    __syncthreads();
    lds_fft_rows<T, +1, EPT, kT>(buf, logM, d.nf_in << d.logRB, ldr, s_tw, 1);
  } else {
    const int n_in = (d.nf_in << d.logRB) << d.lognx;
    for (int idx = tid; idx < n_in; idx += kT) {
      const int x = idx & (nx - 1);
      const int lr = idx >> d.lognx;
      const int f = lr >> d.logRB;
      const int row = row0 + (lr & (d.RB - 1));
      bufr[rpad(lr, x, ldr)] =
          row < d.nrows ? op.inr[f * d.ifs + (long long)row * nx + x] : T(0);
    }
    __syncthreads();
  }

  if constexpr (Op::kStoreReal) {
    const int n_out = (d.nf_in << d.logRB) << d.lognx;
    for (int idx = tid; idx < n_out; idx += kT) {
      const int x = idx & (nx - 1);
      const int lr = idx >> d.lognx;
      const int f = lr >> d.logRB;
      const int row = row0 + (lr & (d.RB - 1));
      if (row < d.nrows) op.outr[f * d.ofs + (long long)row * nx + x] = bufr[rpad(lr, x, ldr)];
    }
    return;
  } else {
    op.physical(bufr, row0, ldr);
    __syncthreads();
    lds_fft_rows<T, -1, EPT, kT>(buf, logM, d.nf_out << d.logRB, ldr, s_tw, 1);
    // r2c post-processing, one thread per pair (k, M - k):
    // E = (Z[k] + conj Z[M-k]) / 2, O = -i (Z[k] - conj Z[M-k]) / 2,
    // X[k] = E + W^k O, X[M-k] = conj(E - W^k O); k = 0 also does M/2.
    const int n_out = (d.nf_out << d.logRB) << logH;
    for (int idx = tid; idx < n_out; idx += kT) {
      const int k = idx & (H - 1);
      const int lr = idx >> logH;
      const int f = lr >> d.logRB;
      const int row = row0 + (lr & (d.RB - 1));
      if (row >= d.nrows) continue;
      const cpx<T>* Z = buf + lr * ldr;
      cpx<T>* dst = op.out + f * d.ofs + (long long)row * d.ld_out;
      const T h = T(0.5);
      {
        const cpx<T> Zk = Z[lpad(k)], Zm = conj(Z[lpad((M - k) & (M - 1))]);
        const cpx<T> E = scale(Zk + Zm, h), dd = Zk - Zm;
        const cpx<T> P = s_tw[k] * mk<T>(dd.y * h, -dd.x * h);
        if (k < d.mx_out) dst[k] = scale(E + P, d.scale);
        if (M - k < d.mx_out) dst[M - k] = scale(conj(E - P), d.scale);
      }
      if (k == 0 && H < d.mx_out) {
        const cpx<T> Zk = Z[lpad(H)], Zm = conj(Z[lpad(H)]);
        const cpx<T> E = scale(Zk + Zm, h), dd = Zk - Zm;
        const cpx<T> P = s_tw[H] * mk<T>(dd.y * h, -dd.x * h);
        dst[H] = scale(E + P, d.scale);
      }
    }
  }
}

// Row pass of any length nx (grids that are not powers of two): each real
// row is one length-nx complex transform in LDS (gen_fft, row layout with
// stride nx + 1).  The c2r input is the half spectrum extended by Hermitian
// symmetry, X[nx - k] = conj X[k]; the real part of the inverse transform is
// the row (the imaginary parts of X[0] and, for even nx, X[nx/2] drop out,
// as in numpy's irfft).  The products run per point (Op::point) on the real
// parts; the forward transform of the real product rows keeps kx < mx_out.
// Rows per workgroup: d.RB (any count), fields stacked as rows f RB + r.
template <typename T, int kT, class Op>
static __global__ __launch_bounds__(kT) void k_xpass_gen(Op op) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const XDesc<T>& d = op.d;
  const int nx = d.nx, ld = nx + 1, RB = d.RB;
  cpx<T>* s_tw = reinterpret_cast<cpx<T>*>(smem_raw);
  const int nfm = d.nf_in > d.nf_out ? d.nf_in : d.nf_out;
  cpx<T>* A = s_tw + nx;
  cpx<T>* B = A + nfm * RB * ld;
  const int row0 = blockIdx.x * RB;
  const int tid = threadIdx.x;
  const TileLayout lay{-1, ld};
  for (int i = tid; i < nx; i += kT) s_tw[i] = d.tw[i];
  const int n_in = d.nf_in * RB * nx;
  for (int idx = tid; idx < n_in; idx += kT) {
    const int lr = idx / nx, x = idx - lr * nx;
    const int f = lr / RB, row = row0 + (lr - f * RB);
    cpx<T> v = zero<T>();
    if (row < d.nrows) {
      if constexpr (Op::kInv) {
        const int k = x <= nx / 2 ? x : nx - x;
        if (k < d.mx_in) v = op.in[f * d.ifs + (long long)row * d.ld_in + k];
        if (x > nx / 2) v = conj(v);
      } else {
        v.x = op.inr[f * d.ifs + (long long)row * nx + x];
      }
    }
    A[lr * ld + x] = v;
  }
  __syncthreads();
  cpx<T>* res = A;
  if constexpr (Op::kInv) {
    res = gen_fft<T, +1, kT>(A, B, lay, s_tw, nx, d.nf_in * RB, d.nst, d.rad);
    if constexpr (Op::kStoreReal) {
      for (int idx = tid; idx < n_in; idx += kT) {
        const int lr = idx / nx, x = idx - lr * nx;
        const int f = lr / RB, row = row0 + (lr - f * RB);
        if (row < d.nrows) op.outr[f * d.ofs + (long long)row * nx + x] = res[lr * ld + x].x;
      }
      return;
    } else {
      T m[3] = {0, 0, 0};
      for (int idx = tid; idx < RB * nx; idx += kT) {
        const int r = idx / nx, x = idx - r * nx;
        T v[8];
        for (int c = 0; c < d.nf_in && c < 8; ++c) v[c] = res[(c * RB + r) * ld + x].x;
        T mm[3] = {0, 0, 0};
        op.point(v, mm);
        if (row0 + r < d.nrows)
          for (int c = 0; c < 3; ++c) m[c] = fmax(m[c], mm[c]);
        for (int c = 0; c < d.nf_out; ++c) res[(c * RB + r) * ld + x] = mk<T>(v[c], T(0));
      }
      if constexpr (Op::kNMax > 0)
        if (op.maxima) block_max_atomic<T, kT>(m[0], m[1], m[2], Op::kNMax, op.maxima);
      __syncthreads();
    }
  }
  cpx<T>* other = res == A ? B : A;
  res = gen_fft<T, -1, kT>(res, other, lay, s_tw, nx, d.nf_out * RB, d.nst, d.rad);
  const int n_out = d.nf_out * RB * d.mx_out;
  for (int idx = tid; idx < n_out; idx += kT) {
    const int lr = idx / d.mx_out, k = idx - lr * d.mx_out;
    const int f = lr / RB, row = row0 + (lr - f * RB);
    if (row < d.nrows)
      op.out[f * d.ofs + (long long)row * d.ld_out + k] = scale(res[lr * ld + k], d.scale);
  }
}

// plain c2r rows: W2 rows -> X rows
template <typename T>
struct OpXC2R {
  static constexpr bool kInv = true, kStoreReal = true;
  static constexpr int kNMax = 0;
  unsigned long long* maxima = nullptr;
  __device__ void point(T*, T*) const {}
  XDesc<T> d;
  const cpx<T>* in;
  const T* inr;
  T* outr;
  cpx<T>* out;
  __device__ void physical(T*, int, int) const {}
};

// plain r2c rows: X rows -> W2 rows, scaled by 1/(number of points)
template <typename T>
struct OpXR2C {
  static constexpr bool kInv = false, kStoreReal = false;
  static constexpr int kNMax = 0;
  unsigned long long* maxima = nullptr;
  __device__ void point(T*, T*) const {}
  XDesc<T> d;
  const cpx<T>* in;
  const T* inr;
  T* outr;
  cpx<T>* out;
  __device__ void physical(T*, int, int) const {}
};

// ns3d rows: (v, omega) -> v x omega, optional CFL maxima of |v_i|.
template <typename T, int kT>
struct OpXNS3D {
  static constexpr bool kInv = true, kStoreReal = false;
  XDesc<T> d;
  const cpx<T>* in;
  const T* inr;
  T* outr;
  cpx<T>* out;
  unsigned long long* maxima;  // 3 slots or null
  static constexpr int kNMax = 3;
  // one point: v = (v, omega) -> v x omega, maxima of |v_i| into m
  __device__ void point(T* v, T* m) const {
    const T vx = v[0], vy = v[1], vz = v[2];
    v[0] = vy * v[5] - vz * v[4];
    v[1] = vz * v[3] - vx * v[5];
    v[2] = vx * v[4] - vy * v[3];
    m[0] = fmax(m[0], fabs(vx));
    m[1] = fmax(m[1], fabs(vy));
    m[2] = fmax(m[2], fabs(vz));
  }
  __device__ void physical(T* b, int row0, int ldr) const {
    const int n = d.RB << d.lognx;
    const int s = 2 * ldr * d.RB;  // reals per field
    T m0 = 0, m1 = 0, m2 = 0;
    for (int idx = threadIdx.x; idx < n; idx += kT) {
      T* p = b + rpad(idx >> d.lognx, idx & (d.nx - 1), ldr);
      const T vx = p[0], vy = p[s], vz = p[2 * s];
      const T wx = p[3 * s], wy = p[4 * s], wz = p[5 * s];
      p[0] = vy * wz - vz * wy;
      p[s] = vz * wx - vx * wz;
      p[2 * s] = vx * wy - vy * wx;
      if (row0 + (idx >> d.lognx) < d.nrows) {
        m0 = fmax(m0, fabs(vx));
        m1 = fmax(m1, fabs(vy));
        m2 = fmax(m2, fabs(vz));
      }
    }
    if (maxima) block_max_atomic<T, kT>(m0, m1, m2, 3, maxima);
  }
};

// ns3d.strat rows: (v, omega, b) -> (v x omega, v b), CFL maxima of |v_i|.
template <typename T, int kT>
struct OpXStrat {
  static constexpr bool kInv = true, kStoreReal = false;
  XDesc<T> d;
  const cpx<T>* in;
  const T* inr;
  T* outr;
  cpx<T>* out;
  unsigned long long* maxima;  // 3 slots or null
  static constexpr int kNMax = 3;
  // one point: (v, omega, b) -> (v x omega, v b), maxima of |v_i| into m
  __device__ void point(T* v, T* m) const {
    const T vx = v[0], vy = v[1], vz = v[2], bb = v[6];
    v[0] = vy * v[5] - vz * v[4];
    v[1] = vz * v[3] - vx * v[5];
    v[2] = vx * v[4] - vy * v[3];
    v[3] = vx * bb;
    v[4] = vy * bb;
    v[5] = vz * bb;
    m[0] = fmax(m[0], fabs(vx));
    m[1] = fmax(m[1], fabs(vy));
    m[2] = fmax(m[2], fabs(vz));
  }
  __device__ void physical(T* b, int row0, int ldr) const {
    const int n = d.RB << d.lognx;
    const int s = 2 * ldr * d.RB;
    T m0 = 0, m1 = 0, m2 = 0;
    for (int idx = threadIdx.x; idx < n; idx += kT) {
      T* p = b + rpad(idx >> d.lognx, idx & (d.nx - 1), ldr);
      const T vx = p[0], vy = p[s], vz = p[2 * s];
      const T wx = p[3 * s], wy = p[4 * s], wz = p[5 * s], bb = p[6 * s];
      p[0] = vy * wz - vz * wy;
      p[s] = vz * wx - vx * wz;
      p[2 * s] = vx * wy - vy * wx;
      p[3 * s] = vx * bb;
      p[4 * s] = vy * bb;
      p[5 * s] = vz * bb;
      if (row0 + (idx >> d.lognx) < d.nrows) {
        m0 = fmax(m0, fabs(vx));
        m1 = fmax(m1, fabs(vy));
        m2 = fmax(m2, fabs(vz));
      }
    }
    if (maxima) block_max_atomic<T, kT>(m0, m1, m2, 3, maxima);
  }
};

// ns2d rows: (ux, uy, d_x rot, d_y rot) -> -ux d_x rot - uy (d_y rot + beta)
template <typename T, int kT>
struct OpXNS2D {
  static constexpr bool kInv = true, kStoreReal = false;
  XDesc<T> d;
  const cpx<T>* in;
  const T* inr;
  T* outr;
  cpx<T>* out;
  unsigned long long* maxima;  // 2 slots or null
  T beta;
  static constexpr int kNMax = 2;
  // one point: (ux, uy, d_x rot, d_y rot) -> -ux d_x rot - uy (d_y rot + beta)
  __device__ void point(T* v, T* m) const {
    const T ux = v[0], uy = v[1], px = v[2], py = v[3];
    v[0] = beta == T(0) ? -ux * px - uy * py : -ux * px - uy * (py + beta);
    m[0] = fmax(m[0], fabs(ux));
    m[1] = fmax(m[1], fabs(uy));
  }
  __device__ void physical(T* b, int row0, int ldr) const {
    const int n = d.RB << d.lognx;
    const int s = 2 * ldr * d.RB;
    T m0 = 0, m1 = 0;
    for (int idx = threadIdx.x; idx < n; idx += kT) {
      T* p = b + rpad(idx >> d.lognx, idx & (d.nx - 1), ldr);
      const T ux = p[0], uy = p[s], px = p[2 * s], py = p[3 * s];
      p[0] = beta == T(0) ? -ux * px - uy * py : -ux * px - uy * (py + beta);
      if (row0 + (idx >> d.lognx) < d.nrows) {
        m0 = fmax(m0, fabs(ux));
        m1 = fmax(m1, fabs(uy));
      }
    }
    if (maxima) block_max_atomic<T, kT>(m0, m1, T(0), 2, maxima);
  }
};

}  // namespace fsx
