// Device-side FFT building blocks for the fluidsim_amd engine (gfx950).
//
// Complex numbers are interleaved (re, im) pairs of T (float or double), the
// layout numpy's complex64/complex128 use, so K arrays cross the C-ABI as
// plain pointers.
//
// The transform is a Stockham autosort FFT of power-of-two length N done in
// LDS: a tile of `lines` lines is staged in LDS, every radix-R stage reads its
// R inputs per butterfly into registers, applies twiddles and a hard-coded
// DFT_R, and writes the outputs back in Stockham order, so the result is in
// natural order without a bit-reversal pass.  Stage radices are 8 while at
// least three factors of two remain, then 4 or 2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsx {

constexpr int kThreads = 256;  // threads per workgroup for every pass kernel

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
  // c[k] *= w8^k with w8 = exp(DIR i pi/4)
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

template <typename T, int DIR, int R>
__device__ __forceinline__ void dftR(cpx<T>* v) {
  if constexpr (R == 8)
    dft8<T, DIR>(v);
  else if constexpr (R == 4)
    dft4<T, DIR>(v);
  else
    dft2<T, DIR>(v);
}

// One Stockham stage of radix R on `lines` lines of length N = 2^logN stored
// in LDS at buf[pos * ps + line * ls].  Ns = 2^logNs is the length of the
// sub-transforms already done.  Butterfly j of a line reads positions
// j + r N/R, multiplies input r by W_{Ns R}^{r k} (k = j mod Ns) and writes
// output r to ((j - k) R + k) + r Ns.
//
// Twiddles come from a global table tw[m] = exp(-2 pi i m / (N << twshift)).
// Each thread owns EPT/R butterflies; the whole tile must hold at most
// EPT * kThreads elements.  line_fast selects the thread -> butterfly map:
// consecutive threads take consecutive lines (strided passes, lines are
// contiguous in LDS) or consecutive butterflies (row passes).
template <typename T, int DIR, int EPT, int R>
__device__ __forceinline__ void fft_stage(cpx<T>* buf, int logN, int logNs,
                                          int lines, bool line_fast, int ps,
                                          int ls, const cpx<T>* __restrict__ tw,
                                          int twshift) {
  constexpr int LOGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  constexpr int NB = EPT / R;
  const int lognb = logN - LOGR;
  const int nb = 1 << lognb;
  const int total = nb * lines;
  const int Ns = 1 << logNs;
  const int tid = threadIdx.x;
  cpx<T> v[NB][R];
  int obase[NB];
  bool act[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = tid + i * kThreads;
    act[i] = b < total;
    if (act[i]) {
      int j, line;
      if (line_fast) {
        line = b % lines;
        j = b / lines;
      } else {
        j = b & (nb - 1);
        line = b >> lognb;
      }
      const int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * nb) * ps + line * ls];
      if (logNs > 0) {
        const int sh = (logN - logNs - LOGR) + twshift;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> w = tw[(r * k) << sh];
          if (DIR > 0) w.y = -w.y;
          v[i][r] = v[i][r] * w;
        }
      }
      dftR<T, DIR, R>(v[i]);
      obase[i] = (((j - k) << LOGR) + k) * ps + line * ls;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (act[i]) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[obase[i] + r * Ns * ps] = v[i][r];
    }
  }
  __syncthreads();
}

// Full in-LDS FFT (all stages).  Caller synchronises before the call.
template <typename T, int DIR, int EPT>
__device__ __forceinline__ void lds_fft(cpx<T>* buf, int logN, int lines,
                                        bool line_fast, int ps, int ls,
                                        const cpx<T>* __restrict__ tw,
                                        int twshift) {
  int logNs = 0;
  while (logNs < logN) {
    const int rem = logN - logNs;
    if (rem >= 3) {
      fft_stage<T, DIR, EPT, 8>(buf, logN, logNs, lines, line_fast, ps, ls, tw,
                                twshift);
      logNs += 3;
    } else if (rem == 2) {
      fft_stage<T, DIR, EPT, 4>(buf, logN, logNs, lines, line_fast, ps, ls, tw,
                                twshift);
      logNs += 2;
    } else {
      fft_stage<T, DIR, EPT, 2>(buf, logN, logNs, lines, line_fast, ps, ls, tw,
                                twshift);
      logNs += 1;
    }
  }
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
