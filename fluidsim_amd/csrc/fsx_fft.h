// Device-side FFT building blocks for the fluidsim_amd engine (gfx950).
//
// Complex numbers are interleaved (re, im) pairs of T (float or double), the
// layout numpy's complex64/complex128 use, so K arrays cross the C-ABI as
// plain pointers.
//
// Transforms are Stockham autosort FFTs of power-of-two length N: radix-R
// stages (R = 8 while at least three factors of two remain, then 4 or 2).
// Butterfly j of a stage with sub-transform length Ns reads positions
// j + r N/R, multiplies input r by W_{Ns R}^{r k} (k = j mod Ns), applies a
// hard-coded DFT_R and writes output r to (j - k) R + k + r Ns, so the
// result is in natural order without a bit-reversal pass.
//
// The passes keep the first stage in registers (inputs come straight from
// HBM) and the last stage in registers (outputs go straight to HBM: for the
// last stage k = j, so output r lands at position j + r Ns); only the middle
// stages exchange data through LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fsx {

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T a, T b) {
  cpx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) {
  return mk<T>(a.x * s, a.y * s);
}
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) {
  return mk<T>(a.x, -a.y);
}
// multiply by i
template <typename T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
  return mk<T>(-a.y, a.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> zero() {
  return mk<T>(T(0), T(0));
}

// a b + c with one rounding, whatever the fp contraction setting
template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
  if constexpr (sizeof(T) == 4)
    return __builtin_fmaf(a, b, c);
  else
    return __builtin_fma(a, b, c);
}

// DIR = -1: forward transform, kernel exp(-2 pi i jk/N).
// DIR = +1: inverse transform, kernel exp(+2 pi i jk/N).
// Multiplication by exp(DIR * i pi / 2): forward -> -i, inverse -> +i.
template <typename T, int DIR>
__device__ __forceinline__ cpx<T> rot90(cpx<T> a) {
  return DIR < 0 ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x);
}

template <typename T, int DIR>
__device__ __forceinline__ void dft2(cpx<T>* v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <typename T, int DIR>
__device__ __forceinline__ void dft4(cpx<T>* v) {
  cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
  cpx<T> t2 = v[1] + v[3], t3 = rot90<T, DIR>(v[1] - v[3]);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}

// 8-point DFT as one radix-2 step (stride 4) followed by two 4-point DFTs:
// X[2m] = DFT4(a[k] + a[k+4]), X[2m+1] = DFT4((a[k] - a[k+4]) w8^k).
template <typename T, int DIR>
__device__ __forceinline__ void dft8(cpx<T>* v) {
  const T h = T(0.70710678118654752440084436210484903928);
  cpx<T> b[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    b[k] = v[k] + v[k + 4];
    c[k] = v[k] - v[k + 4];
  }
  {
    cpx<T> a = c[1];
    c[1] = DIR < 0 ? mk<T>((a.x + a.y) * h, (a.y - a.x) * h)
                   : mk<T>((a.x - a.y) * h, (a.x + a.y) * h);
    c[2] = rot90<T, DIR>(c[2]);
    a = c[3];
    c[3] = DIR < 0 ? mk<T>((a.y - a.x) * h, -(a.x + a.y) * h)
                   : mk<T>(-(a.x + a.y) * h, (a.x - a.y) * h);
  }
  dft4<T, DIR>(b);
  dft4<T, DIR>(c);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    v[2 * m] = b[m];
    v[2 * m + 1] = c[m];
  }
}

// a exp(DIR 2 pi i m / 16) for a constant m
template <typename T, int DIR, int M>
__device__ __forceinline__ cpx<T> w16(cpx<T> a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (m == 4) {
    return rot90<T, DIR>(a);
  } else {
    // cos and sin of 2 pi m / 16 from the first-octant values
    constexpr T c1 = T(0.92387953251128675612818318939678828682);  // cos(pi/8)
    constexpr T s1 = T(0.38268343236508977172845998403039886676);  // sin(pi/8)
    constexpr T h = T(0.70710678118654752440084436210484903928);
    constexpr T ct[16] = {T(1), c1, h, s1, T(0), -s1, -h, -c1, T(-1), -c1, -h, -s1, T(0), s1, h, c1};
    constexpr T c = ct[m], sn = ct[(m + 12) & 15];  // sin x = cos(x - pi/2)
    const T wi = DIR < 0 ? -sn : sn;
    // explicit fused multiply-adds: the rounding is then the same in every
    // pass that inlines this (no choice left to fp contraction)
    return mk<T>(fma_t(a.x, c, -(a.y * wi)), fma_t(a.x, wi, a.y * c));
  }
}

// 16-point DFT as 4 x 4: with n = 4 n1 + n2 and k = k1 + 4 k2,
// X[k1 + 4 k2] = DFT4_{n2}( W16^{n2 k1} DFT4_{n1}(x[4 n1 + n2])[k1] )[k2].
template <typename T, int DIR>
__device__ __forceinline__ void dft16(cpx<T>* v) {
  cpx<T> y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) y[n2][n1] = v[4 * n1 + n2];
    dft4<T, DIR>(y[n2]);
  }
  y[1][1] = w16<T, DIR, 1>(y[1][1]);
  y[1][2] = w16<T, DIR, 2>(y[1][2]);
  y[1][3] = w16<T, DIR, 3>(y[1][3]);
  y[2][1] = w16<T, DIR, 2>(y[2][1]);
  y[2][2] = w16<T, DIR, 4>(y[2][2]);
  y[2][3] = w16<T, DIR, 6>(y[2][3]);
  y[3][1] = w16<T, DIR, 3>(y[3][1]);
  y[3][2] = w16<T, DIR, 6>(y[3][2]);
  y[3][3] = w16<T, DIR, 9>(y[3][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<T> b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<T, DIR>(b);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
  }
}

template <typename T, int DIR, int R>
__device__ __forceinline__ void dftR(cpx<T>* v) {
  static_assert(R == 16 || R == 8 || R == 4 || R == 2, "radix 2, 4, 8 or 16");
  if constexpr (R == 16)
    dft16<T, DIR>(v);
  else if constexpr (R == 8)
    dft8<T, DIR>(v);
  else if constexpr (R == 4)
    dft4<T, DIR>(v);
  else
    dft2<T, DIR>(v);
}

// 3-point DFT: y0 = a + b + c, y1,2 = a - (b + c)/2 +- i DIR (sqrt 3 / 2)(b - c).
template <typename T, int DIR>
__device__ __forceinline__ void dft3(cpx<T>* v) {
  const T h = T(0.86602540378443864676372317075293618347);
  const cpx<T> s = v[1] + v[2];
  const cpx<T> t = v[0] - scale(s, T(0.5));
  const cpx<T> u = rot90<T, DIR>(scale(v[1] - v[2], h));
  v[0] = v[0] + s;
  v[1] = t + u;
  v[2] = t - u;
}

// 5-point DFT from the cosines and sines of 2 pi / 5 and 4 pi / 5:
// y1,4 = x0 + c1 a1 + c2 a2 +- i DIR (s1 b1 + s2 b2),
// y2,3 = x0 + c2 a1 + c1 a2 +- i DIR (s2 b1 - s1 b2),
// a1 = x1 + x4, a2 = x2 + x3, b1 = x1 - x4, b2 = x2 - x3.
template <typename T, int DIR>
__device__ __forceinline__ void dft5(cpx<T>* v) {
  const T c1 = T(0.30901699437494742410229341718281905886);
  const T c2 = T(-0.80901699437494742410229341718281905886);
  const T s1 = T(0.95105651629515357211643933337938214340);
  const T s2 = T(0.58778525229247312916870595463907276860);
  const cpx<T> a1 = v[1] + v[4], a2 = v[2] + v[3];
  const cpx<T> b1 = v[1] - v[4], b2 = v[2] - v[3];
  const cpx<T> t1 = v[0] + scale(a1, c1) + scale(a2, c2);
  const cpx<T> t2 = v[0] + scale(a1, c2) + scale(a2, c1);
  const cpx<T> u1 = rot90<T, DIR>(scale(b1, s1) + scale(b2, s2));
  const cpx<T> u2 = rot90<T, DIR>(scale(b1, s2) - scale(b2, s1));
  v[0] = v[0] + a1 + a2;
  v[1] = t1 + u1;
  v[4] = t1 - u1;
  v[2] = t2 + u2;
  v[3] = t2 - u2;
}

// Radix of stage s of a length-2^logN transform (8s first, then 4 or 2).
__host__ __device__ __forceinline__ int stage_log_radix(int logN, int logNs) {
  const int rem = logN - logNs;
  return rem >= 3 ? 3 : rem;
}

// LDS address with one pad slot per 32 elements: power-of-two strides of
// the Stockham write patterns then fall on different banks.
__device__ __forceinline__ int lpad(int a) { return a + (a >> 5); }

// Apply the Stockham twiddles W_{Ns R}^{r k} to v[1..R-1].
// tw is a table exp(-2 pi i m / Ntab) with Ntab = N << twshift.
template <typename T, int DIR, int R, int LOGR>
__device__ __forceinline__ void twiddle(cpx<T>* v, int k, int logN, int logNs,
                                        const cpx<T>* tw, int twshift) {
  const int sh = (logN - logNs - LOGR) + twshift;
#pragma unroll
  for (int r = 1; r < R; ++r) {
    cpx<T> w = tw[(r * k) << sh];
    if (DIR > 0) w.y = -w.y;
    v[r] = v[r] * w;
  }
}

// One middle Stockham stage in LDS over `lines` lines stored at
// buf[(pos << logL) + line] (line-fast: lines are contiguous) with
// 2^logL lines.  Each of the kT threads owns EPT/R butterflies.
template <typename T, int DIR, int EPT, int R, int kT>
__device__ __forceinline__ void lds_stage_lf(cpx<T>* buf, int logN, int logNs,
                                             int logL, const cpx<T>* tw) {
  constexpr int LOGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  constexpr int NB = EPT / R;
  const int lognb = logN - LOGR;
  const int total = 1 << (lognb + logL);
  const int Ns = 1 << logNs;
  cpx<T> v[NB][R];
  int ob[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * kT;
    if (b < total) {
      const int line = b & ((1 << logL) - 1);
      const int j = b >> logL;
      const int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[i][r] = buf[lpad(((j + (r << lognb)) << logL) + line)];
      twiddle<T, DIR, R, LOGR>(v[i], k, logN, logNs, tw, 0);
      dftR<T, DIR, R>(v[i]);
      ob[i] = ((((j - k) << LOGR) + k) << logL) + line;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * kT;
    if (b < total) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lpad(ob[i] + ((r * Ns) << logL))] = v[i][r];
    }
  }
  __syncthreads();
}

// Full in-LDS FFT over `lines` rows stored at buf[line * ldr + lpad(pos)]
// (row layout, butterfly-fast thread map), all stages.  Used by the row
// (x) pass.  Caller synchronises before the call.
template <typename T, int DIR, int EPT, int R, int kT>
__device__ __forceinline__ void lds_stage_rows(cpx<T>* buf, int logN,
                                               int logNs, int lines, int ldr,
                                               const cpx<T>* tw, int twshift) {
  constexpr int LOGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  constexpr int NB = EPT / R;
  const int lognb = logN - LOGR;
  const int total = lines << lognb;
  const int Ns = 1 << logNs;
  cpx<T> v[NB][R];
  int ob[NB], op[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * kT;
    if (b < total) {
      const int j = b & ((1 << lognb) - 1);
      const int line = b >> lognb;
      const int k = j & (Ns - 1);
      const cpx<T>* row = buf + line * ldr;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = row[lpad(j + (r << lognb))];
      if (logNs > 0) twiddle<T, DIR, R, LOGR>(v[i], k, logN, logNs, tw, twshift);
      dftR<T, DIR, R>(v[i]);
      ob[i] = line * ldr;
      op[i] = ((j - k) << LOGR) + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * kT;
    if (b < total) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[ob[i] + lpad(op[i] + r * Ns)] = v[i][r];
    }
  }
  __syncthreads();
}

template <typename T, int DIR, int EPT, int kT>
__device__ __forceinline__ void lds_fft_rows(cpx<T>* buf, int logN, int lines,
                                             int ldr, const cpx<T>* tw,
                                             int twshift) {
  int logNs = 0;
  while (logNs < logN) {
    const int lr = stage_log_radix(logN, logNs);
    if (lr == 3)
      lds_stage_rows<T, DIR, EPT, 8, kT>(buf, logN, logNs, lines, ldr, tw, twshift);
    else if (lr == 2)
      lds_stage_rows<T, DIR, EPT, 4, kT>(buf, logN, logNs, lines, ldr, tw, twshift);
    else
      lds_stage_rows<T, DIR, EPT, 2, kT>(buf, logN, logNs, lines, ldr, tw, twshift);
    logNs += lr;
  }
}

// ------------------------------------------------------ mixed-radix lengths
// Transforms of any length N (grids that are not powers of two) run as
// Stockham stages of radix 8, 4, 2, 3, 5 and, for any other prime factor p,
// a direct p-point DFT.  Every stage reads one LDS buffer and writes the
// other (no register stash across the barrier), with the same index rule as
// the power-of-two stages: butterfly j (k = j mod Ns) reads positions
// j + r N/R, applies W_{Ns R}^{r k} and DFT_R, and writes output r to
// (j - k) R + k + r Ns.  All twiddles are entries of the length-N table
// tw[m] = exp(-2 pi i m / N).
constexpr int kMaxStages = 16;

// Element (line, pos) of a tile: line-fast (strided passes, `lines` a power
// of two) or row layout (row passes, row stride ld).
struct TileLayout {
  int logL;  // line-fast: pos << logL | line
  int ld;    // row layout (logL < 0): line * ld + pos
  __device__ __forceinline__ int at(int line, int pos) const {
    return logL >= 0 ? lpad((pos << logL) + line) : line * ld + pos;
  }
};

template <typename T, int DIR, int R>
__device__ __forceinline__ void gen_butterfly(const cpx<T>* A, cpx<T>* B,
                                              const TileLayout& lay, const cpx<T>* tw,
                                              int N, int Ns, int line, int j) {
  const int NR = N / R;
  const int k = j % Ns;
  const int step = NR / Ns;  // N / (Ns R)
  cpx<T> v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    v[r] = A[lay.at(line, j + r * NR)];
    if (r > 0 && k > 0) {
      cpx<T> w = tw[r * k * step];
      if (DIR > 0) w.y = -w.y;
      v[r] = v[r] * w;
    }
  }
  if constexpr (R == 3)
    dft3<T, DIR>(v);
  else if constexpr (R == 5)
    dft5<T, DIR>(v);
  else
    dftR<T, DIR, R>(v);
  const int o = (j - k) * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) B[lay.at(line, o + r * Ns)] = v[r];
}

// Direct DFT of a prime radix R without a dedicated butterfly: output r of
// butterfly j is sum_m (x_m W_{Ns R}^{m k}) W_R^{m r}.
template <typename T, int DIR>
__device__ __forceinline__ void gen_direct(const cpx<T>* A, cpx<T>* B, const TileLayout& lay,
                                           const cpx<T>* tw, int N, int R, int Ns,
                                           int line, int j) {
  const int NR = N / R;
  const int k = j % Ns;
  const int step = NR / Ns;
  const int o = (j - k) * R + k;
  for (int r = 0; r < R; ++r) {
    cpx<T> acc = zero<T>();
    for (int m = 0; m < R; ++m) {
      cpx<T> x = A[lay.at(line, j + m * NR)];
      cpx<T> w = tw[m * k * step];
      cpx<T> u = tw[((m * r) % R) * NR];
      if (DIR > 0) {
        w.y = -w.y;
        u.y = -u.y;
      }
      acc = acc + (x * w) * u;
    }
    B[lay.at(line, o + r * Ns)] = acc;
  }
}

// All stages of `lines` transforms of length N over the tile in A (B is the
// second buffer).  The caller synchronises before; returns the buffer that
// holds the result, after a barrier.
template <typename T, int DIR, int kT>
__device__ cpx<T>* gen_fft(cpx<T>* A, cpx<T>* B, const TileLayout& lay, const cpx<T>* tw,
                           int N, int lines, int nst, const int* rad) {
  int Ns = 1;
  for (int s = 0; s < nst; ++s) {
    const int R = rad[s];
    const int nb = (N / R) * lines;
    for (int b = threadIdx.x; b < nb; b += kT) {
      int line, j;
      if (lay.logL >= 0) {
        line = b & ((1 << lay.logL) - 1);
        j = b >> lay.logL;
      } else {
        line = b / (N / R);
        j = b - line * (N / R);
      }
      switch (R) {
        case 2: gen_butterfly<T, DIR, 2>(A, B, lay, tw, N, Ns, line, j); break;
        case 3: gen_butterfly<T, DIR, 3>(A, B, lay, tw, N, Ns, line, j); break;
        case 4: gen_butterfly<T, DIR, 4>(A, B, lay, tw, N, Ns, line, j); break;
        case 5: gen_butterfly<T, DIR, 5>(A, B, lay, tw, N, Ns, line, j); break;
        case 8: gen_butterfly<T, DIR, 8>(A, B, lay, tw, N, Ns, line, j); break;
        default: gen_direct<T, DIR>(A, B, lay, tw, N, R, Ns, line, j); break;
      }
    }
    __syncthreads();
    cpx<T>* t = A;
    A = B;
    B = t;
    Ns *= R;
  }
  return A;
}

// Stage radices of a length-n transform (host): 8s, then 4 or 2, then 3s,
// 5s and the remaining primes in increasing order.  Returns the number of
// stages, or -1 when more than kMaxStages would be needed.
inline int gen_factor(int n, int* rad) {
  int ns = 0;
  int p2 = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++p2;
  }
  while (p2 >= 3) {
    if (ns >= kMaxStages) return -1;
    rad[ns++] = 8;
    p2 -= 3;
  }
  if (p2 > 0) {
    if (ns >= kMaxStages) return -1;
    rad[ns++] = 1 << p2;
  }
  for (int p = 3; n > 1; p += 2) {
    while (n % p == 0) {
      if (ns >= kMaxStages) return -1;
      rad[ns++] = p;
      n /= p;
    }
    if ((long long)p * p > n && n > 1) {
      if (ns >= kMaxStages) return -1;
      rad[ns++] = n;
      n = 1;
    }
  }
  return ns;
}

// Kept-index map of a signed (y or z) wavenumber axis: kept indices are
// [0, lo) U [n - hi, n); compact index j in [0, lo + hi).
struct AxisMap {
  int n, lo, hi;
  __host__ __device__ int nk() const { return lo + hi; }
  __host__ __device__ int full(int j) const {
    return j < lo ? j : j + (n - lo - hi);
  }
  __host__ __device__ int compact(int i) const {
    return i < lo ? i : (i >= n - hi ? i - (n - lo - hi) : -1);
  }
};

// Signed wavenumber index with the Nyquist index mapped to +n/2.
__host__ __device__ __forceinline__ int ksigned(int i, int n) {
  return i <= n / 2 ? i : i - n;
}

}  // namespace fsx
