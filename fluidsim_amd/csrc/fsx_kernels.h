// Pass kernels of the fluidsim_amd engine.
//
// Array layouts in HBM (C order, complex = interleaved re/im):
//   X  (physical)     : (nz, ny, nx) real            (2D: nz = 1)
//   K  (spectral)     : (nz, ny, nkx) complex, nkx = nx/2 + 1
//   W1 (after z-pass) : (nz, nyk, mx) complex        (3D only)
//   W2 (after y-pass) : (nz, ny,  mx) complex
// where mx is the number of kept kx indices and nyk the number of kept ky
// indices of the pass (everything for plain transforms, the bounding box of
// the non-dealiased modes for the solver passes, whose pruned modes are zero
// and are neither loaded nor stored).
//
// A strided pass transforms along y or z: a workgroup stages a tile of L
// adjacent lines (consecutive kx, the contiguous index) of length N in LDS,
// so every global access is a contiguous run of L complex numbers.  A row
// pass transforms along x (real <-> half complex through a complex FFT of
// length nx/2) and is where the physical-space products are fused.
#pragma once

#include "fsx_fft.h"

namespace fsx {

template <typename T>
struct StridedDesc {
  int N, logN;    // FFT length
  int L, logL;    // lines per tile (power of two)
  int nouter;     // number of independent planes
  int nlines;     // lines per plane
  int nf;         // fields transformed one after the other
  const cpx<T>* tw;  // exp(-2 pi i m / N), m < N
};

// ---------------------------------------------------------------- strided
template <typename T, int DIR, int EPT, class Op>
__global__ __launch_bounds__(kThreads) void k_strided(Op op) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
  const StridedDesc<T>& d = op.d;
  const int nchunks = (d.nlines + d.L - 1) >> d.logL;
  const int o = blockIdx.x / nchunks;
  const int c = blockIdx.x - o * nchunks;
  const int tid = threadIdx.x;
  const int ntile = d.N << d.logL;
  for (int f = 0; f < d.nf; ++f) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int idx = tid + i * kThreads;
      if (idx < ntile) {
        const int p = idx >> d.logL;
        const int line = (c << d.logL) + (idx & (d.L - 1));
        buf[idx] = line < d.nlines ? op.load(f, o, line, p) : zero<T>();
      }
    }
    __syncthreads();
    lds_fft<T, DIR, EPT>(buf, d.logN, d.L, true, d.L, 1, d.tw, 0);
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int idx = tid + i * kThreads;
      if (idx < ntile) {
        const int p = idx >> d.logL;
        const int line = (c << d.logL) + (idx & (d.L - 1));
        if (line < d.nlines) op.store(f, o, line, p, buf[idx]);
      }
    }
    __syncthreads();
  }
}

// Plain z-pass, inverse: K (nz, ny, nkx) -> W1 (nz, nyk, mx).
// Lines are kx in [0, mx); planes are the kept ky rows.
template <typename T>
struct OpZInv {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int ny, nkx, mx;
  AxisMap ym, zm;
  __device__ cpx<T> load(int f, int o, int l, int p) const {
    if (zm.compact(p) < 0) return zero<T>();
    const int iy = ym.full(o);
    return in[f][((size_t)p * ny + iy) * nkx + l];
  }
  __device__ void store(int f, int o, int l, int p, cpx<T> v) const {
    out[f][((size_t)p * ym.nk() + o) * mx + l] = v;
  }
};

// ns3d z-pass, inverse, with the curl folded into the load:
// fields 0-2 are v, fields 3-5 are omega = i k x v (+ f at k = 0).
template <typename T>
struct OpZInvCurl {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int ny, nz, nkx, mx;
  AxisMap ym, zm;
  T dkx, dky, dkz;
  int has_f;
  T fcor;
  __device__ cpx<T> load(int f, int o, int l, int p) const {
    if (zm.compact(p) < 0) return zero<T>();
    const int iy = ym.full(o);
    const size_t idx = ((size_t)p * ny + iy) * nkx + l;
    if (f < 3) return in[f][idx];
    const T kx = T(l) * dkx;
    const T ky = T(ksigned(iy, ny)) * dky;
    const T kz = T(ksigned(p, nz)) * dkz;
    cpx<T> r;
    if (f == 3) {
      const cpx<T> vy = in[1][idx], vz = in[2][idx];
      r = mul_i(scale(vz, ky) - scale(vy, kz));
    } else if (f == 4) {
      const cpx<T> vx = in[0][idx], vz = in[2][idx];
      r = mul_i(scale(vx, kz) - scale(vz, kx));
    } else {
      const cpx<T> vx = in[0][idx], vy = in[1][idx];
      r = mul_i(scale(vy, kx) - scale(vx, ky));
      if (has_f && idx == 0) r.x += fcor;
    }
    return r;
  }
  __device__ void store(int f, int o, int l, int p, cpx<T> v) const {
    out[f][((size_t)p * ym.nk() + o) * mx + l] = v;
  }
};

// y-pass, inverse: W1 (nz, nyk, mx) -> W2 (nz, ny, mx); planes are z.
template <typename T>
struct OpYInv3 {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int ny, mx;
  AxisMap ym;
  __device__ cpx<T> load(int f, int z, int l, int p) const {
    const int jy = ym.compact(p);
    if (jy < 0) return zero<T>();
    return in[f][((size_t)z * ym.nk() + jy) * mx + l];
  }
  __device__ void store(int f, int z, int l, int p, cpx<T> v) const {
    out[f][((size_t)z * ny + p) * mx + l] = v;
  }
};

// y-pass, forward: W2 (nz, ny, mx) -> W1 (nz, nyk, mx).
template <typename T>
struct OpYFwd3 {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int ny, mx;
  AxisMap ym;
  __device__ cpx<T> load(int f, int z, int l, int p) const {
    return in[f][((size_t)z * ny + p) * mx + l];
  }
  __device__ void store(int f, int z, int l, int p, cpx<T> v) const {
    const int jy = ym.compact(p);
    if (jy >= 0) out[f][((size_t)z * ym.nk() + jy) * mx + l] = v;
  }
};

// z-pass, forward: W1 (nz, nyk, mx) -> K (nz, ny, nkx); only kept modes are
// written, dealiased ones (mask) as zero.
template <typename T>
struct OpZFwd {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int ny, nkx, mx;
  AxisMap ym, zm;
  const unsigned char* mask;
  __device__ cpx<T> load(int f, int jy, int l, int p) const {
    return in[f][((size_t)p * ym.nk() + jy) * mx + l];
  }
  __device__ void store(int f, int jy, int l, int p, cpx<T> v) const {
    if (zm.compact(p) < 0) return;
    const int iy = ym.full(jy);
    const size_t idx = ((size_t)p * ny + iy) * nkx + l;
    if (mask && mask[idx]) v = zero<T>();
    out[f][idx] = v;
  }
};

// 2D y-pass, inverse: K (ny, nkx) -> W2 (ny, mx).
template <typename T>
struct OpYInv2 {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int nkx, mx;
  AxisMap ym;
  __device__ cpx<T> load(int f, int, int l, int p) const {
    if (ym.compact(p) < 0) return zero<T>();
    return in[f][(size_t)p * nkx + l];
  }
  __device__ void store(int f, int, int l, int p, cpx<T> v) const {
    out[f][(size_t)p * mx + l] = v;
  }
};

// ns2d y-pass, inverse, from the vorticity: fields are
// ux = i ky rot / k^2, uy = -i kx rot / k^2, d_x rot = i kx rot,
// d_y rot = i ky rot (k^2 = 1e-14 at k = 0, the K2_not0 convention).
template <typename T>
struct OpYInv2NS {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int ny, nkx, mx;
  AxisMap ym;
  T dkx, dky;
  __device__ cpx<T> load(int f, int, int l, int p) const {
    if (ym.compact(p) < 0) return zero<T>();
    const cpx<T> r = in[0][(size_t)p * nkx + l];
    const T kx = T(l) * dkx;
    const T ky = T(ksigned(p, ny)) * dky;
    if (f == 0 || f == 1) {
      T k2 = kx * kx + ky * ky;
      if (k2 == T(0)) k2 = T(1e-14);
      const T inv = T(1) / k2;
      return f == 0 ? mul_i(scale(r, ky * inv)) : mul_i(scale(r, -kx * inv));
    }
    return mul_i(scale(r, f == 2 ? kx : ky));
  }
  __device__ void store(int f, int, int l, int p, cpx<T> v) const {
    out[f][(size_t)p * mx + l] = v;
  }
};

// 2D y-pass, forward: W2 (ny, mx) -> K (ny, nkx), kept modes only.
template <typename T>
struct OpYFwd2 {
  StridedDesc<T> d;
  const cpx<T>* in[8];
  cpx<T>* out[8];
  int nkx, mx;
  AxisMap ym;
  const unsigned char* mask;
  __device__ cpx<T> load(int f, int, int l, int p) const {
    return in[f][(size_t)p * mx + l];
  }
  __device__ void store(int f, int, int l, int p, cpx<T> v) const {
    if (ym.compact(p) < 0) return;
    const size_t idx = (size_t)p * nkx + l;
    if (mask && mask[idx]) v = zero<T>();
    out[f][idx] = v;
  }
};

// ---------------------------------------------------------------- rows
template <typename T>
struct XDesc {
  int nx, M, logM;  // M = nx / 2 complex points per row
  int RB;           // rows per workgroup
  int nrows;
  int nf_in, nf_out;
  int mx_in, mx_out;  // kept kx on input (c2r) / output (r2c)
  T scale;            // applied to the r2c output
  const cpx<T>* tw;   // exp(-2 pi i m / nx), m < nx
};

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide max of up to 3 values, then one atomicMax per value on the bit
// pattern of a non-negative double (monotonic as unsigned 64-bit).
template <typename T>
__device__ void block_max_atomic(T m0, T m1, T m2, int nv,
                                 unsigned long long* dst) {
  __shared__ T red[3][kThreads / 64];
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
    for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, red[threadIdx.x][w]);
    const double md = (double)m;
    atomicMax(dst + threadIdx.x, (unsigned long long)__double_as_longlong(md));
  }
  __syncthreads();
}

template <typename T, int EPT, class Op>
__global__ __launch_bounds__(kThreads) void k_xpass(Op op) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
  T* bufr = reinterpret_cast<T*>(smem_raw);
  const XDesc<T>& d = op.d;
  const int M = d.M, logM = d.logM, nx = d.nx, RB = d.RB;
  const int row0 = blockIdx.x * RB;
  const int tid = threadIdx.x;

  if constexpr (Op::kInv) {
    // c2r pre-processing: Z[k] = (X[k] + conj X[M-k])
    //                          + i conj(W^k) (X[k] - conj X[M-k]),
    // W = exp(-2 pi i / nx); only the real parts of X[0] and X[M] are used.
    const int n_in = d.nf_in * RB * M;
    for (int idx = tid; idx < n_in; idx += kThreads) {
      const int k = idx & (M - 1);
      const int lr = idx >> logM;
      const int f = lr / RB;
      const int r = lr - f * RB;
      const int row = row0 + r;
      cpx<T> Xk = zero<T>(), Xm = zero<T>();
      if (row < d.nrows) {
        const cpx<T>* src = op.in[f] + (size_t)row * d.mx_in;
        if (k < d.mx_in) Xk = src[k];
        if (M - k < d.mx_in) Xm = src[M - k];
      }
      if (k == 0) {
        Xk.y = T(0);
        Xm.y = T(0);
      }
      const cpx<T> a = Xk + conj(Xm);
      const cpx<T> t = conj(d.tw[k]) * (Xk - conj(Xm));
      buf[idx] = a + mul_i(t);
    }
    __syncthreads();
    lds_fft<T, +1, EPT>(buf, logM, d.nf_in * RB, false, 1, M, d.tw, 1);
  } else {
    // load real rows
    const int n_in = d.nf_in * RB * nx;
    for (int idx = tid; idx < n_in; idx += kThreads) {
      const int x = idx & (nx - 1);
      const int lr = idx / nx;
      const int f = lr / RB;
      const int r = lr - f * RB;
      const int row = row0 + r;
      bufr[idx] = row < d.nrows ? op.in[f][(size_t)row * nx + x] : T(0);
    }
    __syncthreads();
  }

  if constexpr (Op::kStoreReal) {
    const int n_out = d.nf_in * RB * nx;
    for (int idx = tid; idx < n_out; idx += kThreads) {
      const int x = idx & (nx - 1);
      const int lr = idx / nx;
      const int f = lr / RB;
      const int r = lr - f * RB;
      const int row = row0 + r;
      if (row < d.nrows) op.outr[f][(size_t)row * nx + x] = bufr[idx];
    }
    return;
  } else {
    op.physical(bufr, row0);
    __syncthreads();
    lds_fft<T, -1, EPT>(buf, logM, d.nf_out * RB, false, 1, M, d.tw, 1);
    // r2c post-processing: X[k] = E[k] + W^k O[k] with
    // E = (Z[k] + conj Z[M-k]) / 2, O = -i (Z[k] - conj Z[M-k]) / 2.
    const int n_out = d.nf_out * RB * d.mx_out;
    for (int idx = tid; idx < n_out; idx += kThreads) {
      const int k = idx % d.mx_out;
      const int lr = idx / d.mx_out;
      const int f = lr / RB;
      const int r = lr - f * RB;
      const int row = row0 + r;
      if (row >= d.nrows) continue;
      const cpx<T>* Z = buf + lr * M;
      const cpx<T> Zk = Z[k & (M - 1)];
      const cpx<T> Zm = conj(Z[(M - k) & (M - 1)]);
      const cpx<T> E = scale(Zk + Zm, T(0.5));
      const cpx<T> dd = Zk - Zm;
      const cpx<T> O = mk<T>(dd.y * T(0.5), -dd.x * T(0.5));
      const cpx<T> X = scale(E + d.tw[k] * O, d.scale);
      op.out[f][(size_t)row * d.mx_out + k] = X;
    }
  }
}

// plain c2r rows: W2 rows (mx_in) -> X rows
template <typename T>
struct OpXC2R {
  static constexpr bool kInv = true, kStoreReal = true;
  XDesc<T> d;
  const cpx<T>* in[8];
  T* outr[8];
  cpx<T>* out[8];
  __device__ void physical(T*, int) const {}
};

// plain r2c rows: X rows -> W2 rows (mx_out), scaled by 1/(number of points)
template <typename T>
struct OpXR2C {
  static constexpr bool kInv = false, kStoreReal = false;
  XDesc<T> d;
  const T* in[8];
  T* outr[8];
  cpx<T>* out[8];
  __device__ void physical(T*, int) const {}
};

// ns3d rows: (v, omega) -> v x omega, optional CFL maxima of |v_i|.
template <typename T>
struct OpXNS3D {
  static constexpr bool kInv = true, kStoreReal = false;
  XDesc<T> d;
  const cpx<T>* in[8];
  T* outr[8];
  cpx<T>* out[8];
  unsigned long long* maxima;  // 3 slots or null
  __device__ void physical(T* b, int row0) const {
    const int nx = d.nx, RB = d.RB;
    T m0 = 0, m1 = 0, m2 = 0;
    for (int idx = threadIdx.x; idx < RB * nx; idx += kThreads) {
      const int r = idx / nx;
      const int x = idx - r * nx;
      const size_t s = (size_t)RB * nx;
      T* p = b + (size_t)r * nx + x;
      const T vx = p[0], vy = p[s], vz = p[2 * s];
      const T wx = p[3 * s], wy = p[4 * s], wz = p[5 * s];
      p[0] = vy * wz - vz * wy;
      p[s] = vz * wx - vx * wz;
      p[2 * s] = vx * wy - vy * wx;
      if (row0 + r < d.nrows) {
        m0 = fmax(m0, fabs(vx));
        m1 = fmax(m1, fabs(vy));
        m2 = fmax(m2, fabs(vz));
      }
    }
    if (maxima) block_max_atomic<T>(m0, m1, m2, 3, maxima);
  }
};

// ns2d rows: (ux, uy, d_x rot, d_y rot) -> -ux d_x rot - uy (d_y rot + beta)
template <typename T>
struct OpXNS2D {
  static constexpr bool kInv = true, kStoreReal = false;
  XDesc<T> d;
  const cpx<T>* in[8];
  T* outr[8];
  cpx<T>* out[8];
  unsigned long long* maxima;  // 2 slots or null
  T beta;
  __device__ void physical(T* b, int row0) const {
    const int nx = d.nx, RB = d.RB;
    T m0 = 0, m1 = 0;
    for (int idx = threadIdx.x; idx < RB * nx; idx += kThreads) {
      const int r = idx / nx;
      const int x = idx - r * nx;
      const size_t s = (size_t)RB * nx;
      T* p = b + (size_t)r * nx + x;
      const T ux = p[0], uy = p[s], px = p[2 * s], py = p[3 * s];
      p[0] = beta == T(0) ? -ux * px - uy * py : -ux * px - uy * (py + beta);
      if (row0 + r < d.nrows) {
        m0 = fmax(m0, fabs(ux));
        m1 = fmax(m1, fabs(uy));
      }
    }
    if (maxima) block_max_atomic<T>(m0, m1, T(0), 2, maxima);
  }
};

}  // namespace fsx
