// Register-resident row (x) pass of the pseudo-spectral tendencies.
//
// A row of nx = 2M real points is handled as a length-M complex transform
// (z[j] = x[2j] + i x[2j+1]) spread over G = M/P lanes of one wave: lane l
// holds z[l + G q], q < P ("strided" layout).  The Stockham stages run in
// registers (radix P first, then radix <= P); between stages a lane writes
// its outputs to a per-wave LDS row buffer and reads the next stage's
// inputs back, so only the rows of one wave share a buffer and a wave-level
// barrier orders the exchange (no workgroup barrier inside the row loop).
//
// With the first stage reading z[l + G r] and the last stage writing
// output r of butterfly j = l + G t to j + r Ns, a lane starts and ends every
// transform in the strided layout (up to a fixed register permutation), so:
//   * the c2r pre-processing Z[k] <- (X[k], X[M-k]) is computed per element
//     from two global loads (k and its mirror), straight into registers;
//   * the physical-space products are pointwise in registers (the same
//     positions for every field of the row);
//   * the forward transform of the products starts from registers;
//   * the r2c post-processing exchanges the mirror element Z[M-k] through
//     the row buffer once.
// Per output field and row that is two LDS exchanges per inner stage plus
// one mirror exchange, against a full LDS round trip per stage and per
// pre/post/product step in the workgroup-wide row pass (k_xpass).
//
// All twiddles a lane needs depend on its lane index only; they are read
// once from the exp(-2 pi i m / nx) table and kept in registers.
#pragma once

#include "fsx_fft.h"

namespace fsx {

// Stage plan of a length-2^LOGM transform with 2^LOGP points per lane:
// radix 2^LOGP first (clamped to LOGM), then radix 2^min(LOGP, remaining).
template <int LOGM, int LOGP>
struct RowPlan {
  static constexpr int M = 1 << LOGM;
  static constexpr int P = 1 << LOGP;
  static constexpr int G = M / P;
  static_assert(LOGM >= LOGP, "row shorter than the points per lane");
  static_assert(G <= 64, "one row must fit in a wave");
  static constexpr int logr(int s) {
    int done = 0;
    for (int i = 0; i < s; ++i) done += (LOGM - done) < LOGP ? (LOGM - done) : LOGP;
    return (LOGM - done) < LOGP ? (LOGM - done) : LOGP;
  }
  static constexpr int lognsof(int s) {
    int done = 0;
    for (int i = 0; i < s; ++i) done += logr(i);
    return done;
  }
  static constexpr int nstages() {
    int s = 0;
    while (lognsof(s) < LOGM) ++s;
    return s;
  }
  // per-lane twiddles of stages 1.. (stage 0 has Ns = 1: none)
  static constexpr int ntw(int s_end) {
    int n = 0;
    for (int s = 1; s < s_end; ++s) n += (P >> logr(s)) * ((1 << logr(s)) - 1);
    return n;
  }
  static constexpr int NTW = ntw(64);
  // register i after the last stage holds position l + G qpos(i)
  static constexpr int qpos(int i) {
    const int lr = logr(nstages() - 1);
    const int t = i >> lr, r = i & ((1 << lr) - 1);
    return t + r * (P >> lr);
  }
  static constexpr int ireg(int q) {
    for (int i = 0; i < P; ++i)
      if (qpos(i) == q) return i;
    return -1;
  }
};

// Padded LDS position in a row buffer: one slot per 8 (FP32) or 4 (FP64)
// elements, so the Stockham writes of stride 8 spread over the banks.
template <typename T>
constexpr int xr_pad_shift() { return sizeof(T) == 4 ? 3 : 2; }
template <typename T>
__device__ __forceinline__ int xpad(int a) {
  return a + (a >> xr_pad_shift<T>());
}
template <typename T>
__host__ __device__ constexpr int xrow_ld(int M) {
  return M + (M >> xr_pad_shift<T>());
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane twiddles: W^k for the c2r/r2c pairing (k = l + G q) and the Stockham
// twiddles W_{Ns R}^{r k} of stages 1.. (forward sign; conjugated for the
// inverse).  tw = exp(-2 pi i m / (2M)), m < 2M.
template <typename T, int LOGM, int LOGP>
struct LaneTw {
  using PL = RowPlan<LOGM, LOGP>;
  static constexpr int NST = PL::NTW > 0 ? PL::NTW : 1;
  cpx<T> pair[PL::P];
  cpx<T> st[NST];

  __device__ __forceinline__ void load(const cpx<T>* tw, int l) {
#pragma unroll
    for (int q = 0; q < PL::P; ++q) pair[q] = tw[l + PL::G * q];
    if constexpr (PL::nstages() > 1) load_stage<1>(tw, l);
  }
  // stage-S twiddle r of butterfly t of lane l
  template <int S>
  __device__ __forceinline__ cpx<T> get(int l, int t, int r) const {
    constexpr int R = 1 << PL::logr(S);
    return st[PL::ntw(S) + t * (R - 1) + (r - 1)];
  }
  // W_{Ns R}^{r k} = tw[r k 2M / (Ns R)] for the lane's butterflies of
  // stage S, stored from offset ntw(S)
  template <int S>
  __device__ __forceinline__ void load_stage(const cpx<T>* tw, int l) {
    constexpr int lr = PL::logr(S), R = 1 << lr, NB = PL::P >> lr;
    constexpr int lns = PL::lognsof(S);
    constexpr int sh = (LOGM + 1) - (lns + lr);
    constexpr int off = PL::ntw(S);
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int j = l + PL::G * t;
      const int k = j & ((1 << lns) - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) st[off + t * (R - 1) + (r - 1)] = tw[(r * k) << sh];
    }
    if constexpr (S + 1 < PL::nstages()) load_stage<S + 1>(tw, l);
  }
};

// One length-M transform of a lane's P registers, strided layout in
// (register r = position l + G r) and out (register i = position
// l + G qpos(i)).  xb is the row's LDS buffer (xrow_ld(M) elements).
template <typename T, int DIR, int LOGM, int LOGP, int S>
__device__ __forceinline__ void row_stages(cpx<T> (&v)[1 << LOGP], cpx<T>* xb,
                                           int l, const LaneTw<T, LOGM, LOGP>& tw) {
  using PL = RowPlan<LOGM, LOGP>;
  constexpr int P = PL::P, G = PL::G, M = PL::M;
  constexpr int lr = PL::logr(S), R = 1 << lr, NB = P >> lr;
  constexpr int Ns = 1 << PL::lognsof(S);
  if constexpr (S > 0) {
    // exchange: write stage S-1 outputs at their Stockham positions, read
    // this stage's inputs j + r M/R
    constexpr int lrp = PL::logr(S - 1), Rp = 1 << lrp, NBp = P >> lrp;
    constexpr int Nsp = 1 << PL::lognsof(S - 1);
    wave_sync();
#pragma unroll
    for (int t = 0; t < NBp; ++t) {
      const int j = l + G * t;
      const int k = j & (Nsp - 1);
      const int base = ((j - k) << lrp) + k;
#pragma unroll
      for (int r = 0; r < Rp; ++r) xb[xpad<T>(base + r * Nsp)] = v[t * Rp + r];
    }
    wave_sync();
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int j = l + G * t;
#pragma unroll
      for (int r = 0; r < R; ++r) v[t * R + r] = xb[xpad<T>(j + r * (M / R))];
    }
    // twiddles
#pragma unroll
    for (int t = 0; t < NB; ++t) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> w = tw.template get<S>(l, t, r);
        if (DIR > 0) w.y = -w.y;
        v[t * R + r] = v[t * R + r] * w;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NB; ++t) dftR<T, DIR, R>(&v[t * R]);
  if constexpr (S + 1 < PL::nstages())
    row_stages<T, DIR, LOGM, LOGP, S + 1>(v, xb, l, tw);
}

template <typename T, int DIR, int LOGM, int LOGP>
__device__ __forceinline__ void row_fft(cpx<T> (&v)[1 << LOGP], cpx<T>* xb, int l,
                                        const LaneTw<T, LOGM, LOGP>& tw) {
  row_stages<T, DIR, LOGM, LOGP, 0>(v, xb, l, tw);
}

// Descriptor of a fused row pass: rows of nf_in c2r inputs (K rows of
// mx_in kept modes, stride ld_in) -> products -> nf_out r2c outputs.
template <typename T>
struct XRDesc {
  int nrows;
  int mx_in, ld_in, mx_out, ld_out;
  long long ifs, ofs;
  T scale;
  const cpx<T>* in;
  cpx<T>* out;
  const cpx<T>* tw;             // exp(-2 pi i m / nx), m < nx
  unsigned long long* maxima;   // CFL maxima slots or null
};

// Physical-space products of the fused row ops on register arrays.
// Each complex register carries two neighbouring real points.
template <typename T>
struct XRNS3D {
  static constexpr int kIn = 6, kOut = 3, kMax = 3;
  static constexpr bool kTwoPhase = false;
  T beta;  // unused
  template <int P>
  __device__ __forceinline__ void apply(cpx<T> (&f)[kIn][P], T (&m)[kMax],
                                        bool valid) const {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const cpx<T> vx = f[0][i], vy = f[1][i], vz = f[2][i];
      const cpx<T> wx = f[3][i], wy = f[4][i], wz = f[5][i];
      f[0][i] = mk<T>(vy.x * wz.x - vz.x * wy.x, vy.y * wz.y - vz.y * wy.y);
      f[1][i] = mk<T>(vz.x * wx.x - vx.x * wz.x, vz.y * wx.y - vx.y * wz.y);
      f[2][i] = mk<T>(vx.x * wy.x - vy.x * wx.x, vx.y * wy.y - vy.y * wx.y);
      if (valid) {
        m[0] = fmax(m[0], fmax(fabs(vx.x), fabs(vx.y)));
        m[1] = fmax(m[1], fmax(fabs(vy.x), fabs(vy.y)));
        m[2] = fmax(m[2], fmax(fabs(vz.x), fabs(vz.y)));
      }
    }
  }
};

// ns3d.strat: (v, omega, b) -> (v x omega, v b) in two phases, so that at
// most six fields are live per lane (seven at once spill at two waves per
// SIMD): phase A transforms v and omega and writes v x omega over omega
// (outputs 0..2), phase B transforms b and writes v b over v (outputs 3..5).
template <typename T>
struct XRStrat {
  static constexpr int kIn = 7, kOut = 6, kMax = 3;
  static constexpr bool kTwoPhase = true;
  static constexpr int kInA = 6, kOutA = 3, kOutAReg = 3;
  T beta;  // unused
  template <int P>
  __device__ __forceinline__ void apply_a(cpx<T> (&f)[kInA][P], T (&m)[kMax],
                                          bool valid) const {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const cpx<T> vx = f[0][i], vy = f[1][i], vz = f[2][i];
      const cpx<T> wx = f[3][i], wy = f[4][i], wz = f[5][i];
      f[3][i] = mk<T>(vy.x * wz.x - vz.x * wy.x, vy.y * wz.y - vz.y * wy.y);
      f[4][i] = mk<T>(vz.x * wx.x - vx.x * wz.x, vz.y * wx.y - vx.y * wz.y);
      f[5][i] = mk<T>(vx.x * wy.x - vy.x * wx.x, vx.y * wy.y - vy.y * wx.y);
      if (valid) {
        m[0] = fmax(m[0], fmax(fabs(vx.x), fabs(vx.y)));
        m[1] = fmax(m[1], fmax(fabs(vy.x), fabs(vy.y)));
        m[2] = fmax(m[2], fmax(fabs(vz.x), fabs(vz.y)));
      }
    }
  }
  template <int P>
  __device__ __forceinline__ void apply_b(cpx<T> (&f)[kInA][P], const cpx<T> (&b)[P]) const {
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) f[c][i] = mk<T>(f[c][i].x * b[i].x, f[c][i].y * b[i].y);
  }
};

// ns2d: (ux, uy, d_x rot, d_y rot) -> -ux d_x rot - uy (d_y rot + beta)
template <typename T>
struct XRNS2D {
  static constexpr int kIn = 4, kOut = 1, kMax = 2;
  static constexpr bool kTwoPhase = false;
  T beta;
  template <int P>
  __device__ __forceinline__ void apply(cpx<T> (&f)[kIn][P], T (&m)[kMax],
                                        bool valid) const {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const cpx<T> ux = f[0][i], uy = f[1][i], px = f[2][i], py = f[3][i];
      if (beta == T(0))
        f[0][i] = mk<T>(-ux.x * px.x - uy.x * py.x, -ux.y * px.y - uy.y * py.y);
      else
        f[0][i] = mk<T>(-ux.x * px.x - uy.x * (py.x + beta),
                        -ux.y * px.y - uy.y * (py.y + beta));
      if (valid) {
        m[0] = fmax(m[0], fmax(fabs(ux.x), fabs(ux.y)));
        m[1] = fmax(m[1], fmax(fabs(uy.x), fabs(uy.y)));
      }
    }
  }
};

// Buffer-resource loads and stores of one complex value.  A raw buffer
// access whose byte offset is at or beyond the resource's size reads zero
// and writes nothing, so masked elements (dealiased modes, rows past the
// end) take the offset kOOB instead of a branch.
constexpr unsigned kOOB = 0x80000000u;
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base,
                                                          unsigned bytes) {
  // word 3 = 0x00020000: raw (untyped) 32-bit data format on gfx9 parts
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                           (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ cpx<T> buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
template <typename T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned off,
                                          cpx<T> v) {
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), r, off, 0, 0);
}


// Fused row pass, kT threads per workgroup (kT/64 waves, 64/G rows per
// wave), grid-stride over rows.  LDS: one row buffer per row of a wave.
// Global accesses go through one buffer resource per (wave, field) that
// spans the wave's valid rows.
template <typename T, int LOGM, int LOGP, int kT, class Op>
static __global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(2)))
void k_xrow(XRDesc<T> d, Op op) {
  using PL = RowPlan<LOGM, LOGP>;
  constexpr int P = PL::P, G = PL::G, M = PL::M;
  constexpr int RPW = 64 / G;            // rows per wave
  constexpr int RPB = (kT / 64) * RPW;   // rows per workgroup iteration
  constexpr int LD = xrow_ld<T>(M);
  constexpr unsigned cs = sizeof(cpx<T>);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l = lane & (G - 1), rw = lane / G;
  cpx<T>* xb = reinterpret_cast<cpx<T>*>(smem_raw) + (wave * RPW + rw) * LD;

  LaneTw<T, LOGM, LOGP> tw;
  tw.load(d.tw, l);

  // per-lane byte offsets within the wave's rows (kOOB where masked)
  const unsigned rin = (unsigned)rw * d.ld_in * cs, rout = (unsigned)rw * d.ld_out * cs;
  unsigned offk[P], offm[P], offo[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const int k = l + G * q, km = M - k;
    offk[q] = k < d.mx_in ? rin + k * cs : kOOB;
    offm[q] = km < d.mx_in ? rin + km * cs : kOOB;
    offo[q] = k < d.mx_out ? rout + k * cs : kOOB;
  }
  // Nyquist output X[M] (kept only without x dealiasing), lane l = 0
  const unsigned offn = (l == 0 && M < d.mx_out) ? rout + M * cs : kOOB;

  T mloc[Op::kMax];
#pragma unroll
  for (int i = 0; i < Op::kMax; ++i) mloc[i] = T(0);

  // prefetch across row groups where the 32 (FP32) extra live registers
  // fit without spilling (measured on the instantiations: not FP64, not
  // the six-field ns3d rows of length >= 1024)
  constexpr bool kNext = sizeof(T) == 4 && (LOGM < 9 || Op::kIn < 6 || Op::kTwoPhase);
  const int rstride = gridDim.x * RPB;
  // valid rows of the wave in the row group starting at row0 (uniform: kept
  // in scalar registers so the resources need no waterfall)
  auto wave_rows = [&](int row0) {
    int nv = d.nrows - (row0 + wave * RPW);
    nv = nv < 0 ? 0 : (nv > RPW ? RPW : nv);
    return __builtin_amdgcn_readfirstlane(nv);
  };
  // Input fields are loaded into (xk, xm) one transform ahead: the loads of
  // field fi + 1 are issued after the c2r pre-processing of field fi and
  // before its inverse transform, and those of field 0 of the next row
  // group before the last forward transform, so they are in flight during
  // the transforms' exchanges.
  cpx<T> xk[P], xm[P];
  auto fetch_at = [&](int row0, int fi) {
    const __amdgpu_buffer_rsrc_t rs =
        buf_rsrc(d.in + fi * d.ifs + (long long)(row0 + wave * RPW) * d.ld_in,
                 (unsigned)wave_rows(row0) * d.ld_in * cs);
#pragma unroll
    for (int q = 0; q < P; ++q) {
      xk[q] = buf_load<T>(rs, offk[q]);
      xm[q] = buf_load<T>(rs, offm[q]);
    }
  };
  if (kNext && blockIdx.x * RPB < d.nrows) fetch_at(blockIdx.x * RPB, 0);

  for (int row0 = blockIdx.x * RPB; row0 < d.nrows; row0 += rstride) {
    const int wrow0 = row0 + wave * RPW;
    const int nv = wave_rows(row0);
    const bool valid = rw < nv;
    const unsigned out_bytes = (unsigned)nv * d.ld_out * cs;
    const bool more = kNext && row0 + rstride < d.nrows;
    if (!kNext) fetch_at(row0, 0);
    auto fetch = [&](int fi) { fetch_at(row0, fi); };
    // c2r pre-processing of the fetched field into registers, the next
    // field's fetch (next >= 0), then the inverse transform
    auto load_inv = [&](int next, cpx<T>(&z)[P]) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        cpx<T> Xk = xk[q];
        cpx<T> Xm = xm[q];
        if (q == 0 && l == 0) {
          Xk.y = T(0);
          Xm.y = T(0);
        }
        // z = (Xk + conj Xm) + i conj(w) (Xk - conj Xm), as 4 adds + 4 FMAs
        const cpx<T> w = tw.pair[q];
        const T ax = Xk.x + Xm.x, ay = Xk.y - Xm.y;
        const T dx = Xk.x - Xm.x, dy = Xk.y + Xm.y;
        z[q] = mk<T>(fma_t(w.y, dx, fma_t(-w.x, dy, ax)), fma_t(w.y, dy, fma_t(w.x, dx, ay)));
      }
      if (next >= 0) fetch(next);
      row_fft<T, +1, LOGM, LOGP>(z, xb, l, tw);
    };
    // forward transform of product registers z and r2c post-processing into
    // output field fo
    auto fwd_store = [&](int fo, const cpx<T>(&z)[P]) {
      cpx<T> v[P];
#pragma unroll
      for (int r = 0; r < P; ++r) v[r] = z[PL::ireg(r)];
      row_fft<T, -1, LOGM, LOGP>(v, xb, l, tw);
      // mirror exchange: Z[(M - k) mod M] for k = l + G qpos(i)
      wave_sync();
#pragma unroll
      for (int i = 0; i < P; ++i) xb[xpad<T>(l + G * PL::qpos(i))] = v[i];
      wave_sync();
      const __amdgpu_buffer_rsrc_t rd =
          buf_rsrc(d.out + fo * d.ofs + (long long)wrow0 * d.ld_out, out_bytes);
      // X[k] = s (E + P), X[M] = s conj(E - P) with E = (Zk + Zm) / 2,
      // P = w (Zk - Zm) / (2 i), Zm = conj Z[M - k]: with a = Zk + Zm,
      // dd = Zk - Zm and p = w (dd.y, -dd.x), X[k] = (s / 2) (a + p)
      const T hs = T(0.5) * d.scale;
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int q = PL::qpos(i);
        const int k = l + G * q;
        const cpx<T> Zk = v[i];
        const cpx<T> Zm = conj(xb[xpad<T>((M - k) & (M - 1))]);
        const cpx<T> w = tw.pair[q];
        const T ax = Zk.x + Zm.x, ay = Zk.y + Zm.y;
        const T dx = Zk.x - Zm.x, dy = Zk.y - Zm.y;
        if (q == 0) {
          const T px = fma_t(w.x, dy, w.y * dx), py = fma_t(w.y, dy, -(w.x * dx));
          buf_store<T>(rd, offo[q], mk<T>((ax + px) * hs, (ay + py) * hs));
          buf_store<T>(rd, offn, mk<T>((ax - px) * hs, (py - ay) * hs));
        } else {
          buf_store<T>(rd, offo[q], mk<T>(fma_t(w.x, dy, fma_t(w.y, dx, ax)) * hs,
                                          fma_t(w.y, dy, fma_t(-w.x, dx, ay)) * hs));
        }
      }
    };
    if constexpr (Op::kTwoPhase) {
      // phase A from inputs 0..kInA-1, phase B adds input kInA: fewer
      // fields live at once (see the op)
      cpx<T> f[Op::kInA][P];
#pragma unroll
      for (int fi = 0; fi < Op::kInA; ++fi) load_inv(fi + 1 < Op::kInA ? fi + 1 : -1, f[fi]);
      op.template apply_a<P>(f, mloc, valid);
#pragma unroll
      for (int fo = 0; fo < Op::kOutA; ++fo) {
        // input kInA in flight during phase A's last forward transform
        if (fo == Op::kOutA - 1) fetch(Op::kInA);
        fwd_store(fo, f[Op::kOutAReg + fo]);
      }
      cpx<T> b[P];
      load_inv(-1, b);
      op.template apply_b<P>(f, b);
#pragma unroll
      for (int fo = Op::kOutA; fo < Op::kOut; ++fo) {
        if (fo == Op::kOut - 1 && more) fetch_at(row0 + rstride, 0);
        fwd_store(fo, f[fo - Op::kOutA]);
      }
    } else {
      cpx<T> f[Op::kIn][P];
#pragma unroll
      for (int fi = 0; fi < Op::kIn; ++fi) load_inv(fi + 1 < Op::kIn ? fi + 1 : -1, f[fi]);
      op.template apply<P>(f, mloc, valid);
      // forward transforms of the products and r2c post-processing
#pragma unroll
      for (int fo = 0; fo < Op::kOut; ++fo) {
        if (fo == Op::kOut - 1 && more) fetch_at(row0 + rstride, 0);
        fwd_store(fo, f[fo]);
      }
    }
  }

  if (d.maxima) {
    // workgroup max of each CFL component, one atomicMax per component
    __shared__ T red[Op::kMax][kT / 64];
#pragma unroll
    for (int i = 0; i < Op::kMax; ++i) {
      T m = mloc[i];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
      if (lane == 0) red[i][wave] = m;
    }
    __syncthreads();
    if (threadIdx.x < Op::kMax) {
      T m = red[threadIdx.x][0];
      for (int w = 1; w < kT / 64; ++w) m = fmax(m, red[threadIdx.x][w]);
      atomicMax(d.maxima + threadIdx.x,
                (unsigned long long)__double_as_longlong((double)m));
    }
  }
}

}  // namespace fsx
