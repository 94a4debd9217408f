// Pointwise spectral kernels of the fluidsim_amd engine: integrating-factor
// Runge-Kutta stages (with the Leray / toroidal / poloidal projection folded
// in), standalone projection and dealiasing, FP64 spectral sums and the CFL
// time-step update.
#pragma once

#include "fsx_fft.h"

namespace fsx {

// Box of kept modes iterated by the pointwise kernels:
// kz in zm (compact), ky in ym (compact), kx in [0, mx).
struct Box {
  int dim, nx, ny, nz, nkx, mx;
  AxisMap ym, zm;
  double dkx, dky, dkz;
  __device__ size_t count() const {
    return (size_t)(dim == 3 ? zm.nk() : 1) * ym.nk() * mx;
  }
  // compact linear index -> full K index and integer wavenumbers
  __device__ size_t locate(size_t i, int& ikx, int& iky, int& ikz) const {
    ikx = (int)(i % mx);
    size_t rest = i / mx;
    const int jy = (int)(rest % ym.nk());
    const int jz = (int)(rest / ym.nk());
    const int iy = ym.full(jy);
    const int iz = dim == 3 ? zm.full(jz) : 0;
    iky = ksigned(iy, ny);
    ikz = dim == 3 ? ksigned(iz, nz) : 0;
    return ((size_t)iz * ny + iy) * nkx + ikx;
  }
};

struct Visc {
  double nu2, nu4, nu8, num4;
};

// sigma(k) = nu2 k^2 + nu4 k^4 + nu8 k^8 (+ nu_m4 / k^4, whose k = 0 value
// is the one of the first kx mode), fluidsim/base/solvers/pseudo_spect.py
// :134-191.
__device__ __forceinline__ double freq_diss(const Box& b, const Visc& v,
                                            int ikx, int iky, int ikz) {
  const double kx = ikx * b.dkx, ky = iky * b.dky, kz = ikz * b.dkz;
  const double k2 = kx * kx + ky * ky + kz * kz;
  double s = 0.0;
  if (v.nu2 > 0) s = v.nu2 * k2;
  const double k4 = k2 * k2;
  if (v.nu4 > 0) s += v.nu4 * k4;
  if (v.nu8 > 0) s += v.nu8 * (k4 * k4);
  if (v.num4 != 0) {
    double k2n = k2;
    if (ikx == 0 && iky == 0 && ikz == 0) k2n = b.dkx * b.dkx;
    s += v.num4 / (k2n * k2n);
  }
  return s;
}

// projection kinds
enum { PROJ_NONE = 0, PROJ_PERPK = 1, PROJ_TOROIDAL = 2, PROJ_POLOIDAL = 3 };

template <typename T>
__device__ __forceinline__ void project3(cpx<T>& a, cpx<T>& b, cpx<T>& c,
                                         int kind, int ikx, int iky, int ikz,
                                         const Box& bx) {
  const T kx = T(ikx * bx.dkx), ky = T(iky * bx.dky), kz = T(ikz * bx.dkz);
  if (kind == PROJ_PERPK) {
    T k2 = kx * kx + ky * ky + kz * kz;
    if (ikx == 0 && iky == 0 && ikz == 0) k2 = T(1e-14);
    const T inv = T(1) / k2;
    const cpx<T> t = scale(scale(a, kx) + scale(b, ky) + scale(c, kz), inv);
    a = a - scale(t, kx);
    b = b - scale(t, ky);
    c = c - scale(t, kz);
  } else if (kind == PROJ_TOROIDAL || kind == PROJ_POLOIDAL) {
    const T kh2 = kx * kx + ky * ky;
    T kh2n = kh2 == T(0) ? T(1e-14) : kh2;
    T k2n = kh2 + kz * kz;
    if (k2n == T(0)) k2n = T(1e-14);
    const T cphi = kx * sqrt(T(1) / kh2n), sphi = ky * sqrt(T(1) / kh2n);
    if (kind == PROJ_TOROIDAL) {
      const cpx<T> t = scale(b, cphi) - scale(a, sphi);
      a = scale(t, -sphi);
      b = scale(t, cphi);
      c = zero<T>();
    } else {
      const T cth = kz * sqrt(T(1) / k2n), sth = sqrt(kh2 / k2n);
      const cpx<T> t =
          scale(a, cth * cphi) + scale(b, cth * sphi) - scale(c, sth);
      a = scale(t, cth * cphi);
      b = scale(t, cth * sphi);
      c = scale(t, -sth);
    }
  }
}

template <typename T>
struct RKArgs {
  cpx<T>* S[4];
  cpx<T>* tmp[4];
  cpx<T>* t1[4];
  cpx<T>* Nt[4];  // tendencies (projected in place when project != 0)
  const cpx<T>* src[4];  // state the tendency is evaluated on (fused strat)
  int nvar;
  int scheme;  // 4: RK4, 2: RK2
  int stage;    // < 0: store the (projected) tendency into Nt, no update
  int project;  // projection kind applied to variables 0..2 (3D)
  int no_vz_kz0;
  const double* dt;
  Visc visc;
  Box box;
  unsigned int* nanflag;
  // coarse forcing box added to the tendency before the projection (null:
  // no forcing): fc[v * fvs + (cz * fnyc + cy) * fnkxc + cx]
  const cpx<T>* fc;
  long long fvs;
  int fnyc, fnzc, fnkxc;
  // full-size additive tendencies (null: none), added with the forcing
  const cpx<T>* ext[4];
};

// Index in the coarse forcing box of the mode (ikx, iky, ikz) (signed ky,
// kz), or -1 outside: the coarse grid of n_c points holds signed
// wavenumbers -n_c/2 < k <= n_c/2 (put_coarse_array_in_array_fft,
// fluidsim/operators/operators2d.py:525-600).
__device__ __forceinline__ int coarse_index(int ikx, int iky, int ikz, int nkxc,
                                            int nyc, int nzc) {
  if (ikx >= nkxc || iky <= -nyc / 2 || iky > nyc / 2) return -1;
  if (nzc > 1 && (ikz <= -nzc / 2 || ikz > nzc / 2)) return -1;
  const int cy = iky >= 0 ? iky : iky + nyc;
  const int cz = nzc > 1 ? (ikz >= 0 ? ikz : ikz + nzc) : 0;
  return (cz * nyc + cy) * nkxc + ikx;
}

// Integrating-factor Runge-Kutta update of one mode
// (fluidsim/base/time_stepping/pseudo_spect.py:51-68, 466-517, 898-984):
// N holds the NV tendencies of the mode (projected here when a.project),
// E = exp(-dt sigma), E2 = exp(-dt/2 sigma) in FP64.
// (visc: the viscosities of this mode's run; a.visc except in batches)
template <typename T, int NV>
__device__ __forceinline__ void rk_point(const RKArgs<T>& a, const Visc& visc, double dt,
                                         size_t idx, int ikx, int iky, int ikz,
                                         cpx<T>* N) {
  const Box& bx = a.box;
  const double sig = freq_diss(bx, visc, ikx, iky, ikz);
  const T e = T(exp(-dt * sig));
  const T e2 = T(exp(-dt / 2 * sig));
  const T c6 = T(dt / 6), c3 = T(dt / 3), c2 = T(dt / 2), c1 = T(dt);
  if (a.fc) {
    // tendencies += forcing (ns2d/solver.py:190-191, ns3d/solver.py:243-244)
    const int c = coarse_index(ikx, iky, ikz, a.fnkxc, a.fnyc, a.fnzc);
    if (c >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) N[v] = N[v] + a.fc[v * a.fvs + c];
    }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v)
    if (a.ext[v]) N[v] = N[v] + a.ext[v][idx];
  if constexpr (NV >= 3) {
    if (a.project != PROJ_NONE) {
      project3(N[0], N[1], N[2], a.project, ikx, iky, ikz, bx);
      if (a.no_vz_kz0 && ikz == 0) {
        N[2] = zero<T>();
        if (NV > 3) N[3] = zero<T>();
      }
    }
  }
  if (a.stage < 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) a.Nt[v][idx] = N[v];
    return;
  }
  cpx<T> Sn[NV];
  bool last = false;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const cpx<T> Nv = N[v];
    if (a.scheme == 4) {
      if (a.stage == 0) {
        const cpx<T> S = a.S[v][idx];
        a.tmp[v][idx] = scale(S + scale(Nv, c6), e);
        a.t1[v][idx] = scale(S + scale(Nv, c2), e2);
      } else if (a.stage == 1) {
        const cpx<T> S = a.S[v][idx];
        a.tmp[v][idx] = a.tmp[v][idx] + scale(Nv, c3 * e2);
        a.t1[v][idx] = scale(S, e2) + scale(Nv, c2);
      } else if (a.stage == 2) {
        const cpx<T> S = a.S[v][idx];
        a.tmp[v][idx] = a.tmp[v][idx] + scale(Nv, c3 * e2);
        a.t1[v][idx] = scale(S, e) + scale(Nv, c1 * e2);
      } else {
        Sn[v] = a.tmp[v][idx] + scale(Nv, c6);
        last = true;
      }
    } else {
      const cpx<T> S = a.S[v][idx];
      if (a.stage == 0) {
        a.t1[v][idx] = scale(S + scale(Nv, c2), e2);
      } else {
        Sn[v] = scale(S, e) + scale(Nv, c1 * e2);
        last = true;
      }
    }
  }
  if (last) {
    if constexpr (NV >= 3) {
      if (a.project != PROJ_NONE) {
        project3(Sn[0], Sn[1], Sn[2], a.project, ikx, iky, ikz, bx);
        if (a.no_vz_kz0 && ikz == 0) {
          Sn[2] = zero<T>();
          if (NV > 3) Sn[3] = zero<T>();
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) a.S[v][idx] = Sn[v];
    if (isnan(Sn[0].x) || isnan(Sn[0].y)) atomicOr(a.nanflag, 1u);
  }
}

// One Runge-Kutta stage over the kept modes from a tendency array.
// One wavefront per kept (kz, ky) row, lanes along kx.
template <typename T, int NV>
static __global__ __launch_bounds__(256) void k_rk_stage(RKArgs<T> a) {
  const Box& bx = a.box;
  const int nyk = bx.ym.nk();
  const int nrow = (bx.dim == 3 ? bx.zm.nk() : 1) * nyk;
  const int lane = threadIdx.x & 63;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const double dt = *a.dt;
  for (int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < nrow;
       row += nwaves) {
    const int jz = row / nyk, jy = row - jz * nyk;
    const int iy = bx.ym.full(jy);
    const int iz = bx.dim == 3 ? bx.zm.full(jz) : 0;
    const int iky = ksigned(iy, bx.ny);
    const int ikz = bx.dim == 3 ? ksigned(iz, bx.nz) : 0;
    const int base = (iz * bx.ny + iy) * bx.nkx;
    for (int ikx = lane; ikx < bx.mx; ikx += 64) {
      const int idx = base + ikx;
      cpx<T> N[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) N[v] = a.Nt[v][idx];
      rk_point<T, NV>(a, a.visc, dt, idx, ikx, iky, ikz, N);
    }
  }
}

// Standalone projection of 3 fields over the full K array.
template <typename T>
static __global__ __launch_bounds__(256) void k_project(cpx<T>* a, cpx<T>* b,
                                                 cpx<T>* c, int kind, Box box) {
  const size_t n = (size_t)box.nz * box.ny * box.nkx;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ikx = (int)(i % box.nkx);
    const size_t rest = i / box.nkx;
    const int iy = (int)(rest % box.ny), iz = (int)(rest / box.ny);
    cpx<T> va = a[i], vb = b[i], vc = c[i];
    project3(va, vb, vc, kind, ikx, ksigned(iy, box.ny),
             box.dim == 3 ? ksigned(iz, box.nz) : 0, box);
    a[i] = va;
    b[i] = vb;
    c[i] = vc;
  }
}

// Zero the dealiased modes of nf fields (mask or index box).
template <typename T>
static __global__ __launch_bounds__(256) void k_dealias(cpx<T>* base, size_t fstride,
                                                 int nf, Box box,
                                                 const unsigned char* mask) {
  const size_t n = (size_t)box.nz * box.ny * box.nkx;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    bool dealiased;
    if (mask) {
      dealiased = mask[i] != 0;
    } else {
      const int ikx = (int)(i % box.nkx);
      const size_t rest = i / box.nkx;
      const int iy = (int)(rest % box.ny), iz = (int)(rest / box.ny);
      dealiased = ikx >= box.mx || box.ym.compact(iy) < 0 ||
                  (box.dim == 3 && box.zm.compact(iz) < 0);
    }
    if (dealiased)
      for (int f = 0; f < nf; ++f) base[f * fstride + i] = zero<T>();
  }
}

// FP64 Parseval-weighted sums over the full K array.
//   kind 0 (ns2d, field = rot): E, Z, epsK, epsK_hypo, epsZ, epsZ_hypo
//   kind 1 (3D, fields = v):    E, epsK, epsK_hypo
//   kind 2 (strat, v and b):    E_K, E_A (= |b|^2 / (2 N^2)), epsK, epsA
template <typename T>
static __global__ __launch_bounds__(256) void k_sums(const cpx<T>* f0,
                                              const cpx<T>* f1,
                                              const cpx<T>* f2,
                                              const cpx<T>* f3, int kind,
                                              double N2, Box box, Visc visc,
                                              double* out) {
  const size_t n = (size_t)box.nz * box.ny * box.nkx;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const bool even = (box.nx & 1) == 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ikx = (int)(i % box.nkx);
    const size_t rest = i / box.nkx;
    const int iy = (int)(rest % box.ny), iz = (int)(rest / box.ny);
    const int iky = ksigned(iy, box.ny);
    const int ikz = box.dim == 3 ? ksigned(iz, box.nz) : 0;
    const double w =
        (ikx == 0 || (even && ikx == box.nx / 2)) ? 1.0 : 2.0;
    const double kx = ikx * box.dkx, ky = iky * box.dky, kz = ikz * box.dkz;
    const double k2 = kx * kx + ky * ky + kz * kz;
    // f_d (small scales) and f_d_hypo separately
    Visc vs = visc;
    vs.num4 = 0.0;
    const double fd = freq_diss(box, vs, ikx, iky, ikz);
    Visc vh = {0.0, 0.0, 0.0, visc.num4};
    const double fh = visc.num4 != 0 ? freq_diss(box, vh, ikx, iky, ikz) : 0.0;
    if (kind == 0) {
      const cpx<T> r = f0[i];
      const double a2 = (double)r.x * r.x + (double)r.y * r.y;
      const double k2n = k2 == 0 ? 1e-14 : k2;
      const double e = 0.5 * a2 * k2 / (k2n * k2n);
      const double z = 0.5 * a2;
      acc[0] += w * e;
      acc[1] += w * z;
      acc[2] += w * fd * 2 * e;
      acc[3] += w * fh * 2 * e;
      acc[4] += w * fd * 2 * z;
      acc[5] += w * fh * 2 * z;
    } else {
      const cpx<T> a = f0[i], b = f1[i], c = f2[i];
      const double e =
          0.5 * ((double)a.x * a.x + (double)a.y * a.y + (double)b.x * b.x +
                 (double)b.y * b.y + (double)c.x * c.x + (double)c.y * c.y);
      acc[0] += w * e;
      if (kind == 1) {
        acc[1] += w * fd * 2 * e;
        acc[2] += w * fh * 2 * e;
      } else {
        const cpx<T> bb = f3[i];
        const double ea = 0.5 * ((double)bb.x * bb.x + (double)bb.y * bb.y) / N2;
        acc[1] += w * ea;
        acc[2] += w * (fd + fh) * 2 * e;
        acc[3] += w * (fd + fh) * 2 * ea;
      }
    }
  }
  __shared__ double red[6][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < 6; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[q][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0;
    for (int w2 = 0; w2 < 4; ++w2) s += red[threadIdx.x][w2];
    atomicAdd(out + threadIdx.x, s);
  }
}

// CFL update (fluidsim/base/time_stepping/base.py:320-366): tmp =
// sum_i max|u_i| / dx_i, dt_CFL = CFL / tmp, new = min(dt_CFL, dt_max),
// kept unless the relative change exceeds 2 %.
static __global__ void k_cfl_update(const unsigned long long* maxima, int ncomp,
                             double dx, double dy, double dz, double cfl,
                             double dt_max, double* dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double d[3] = {dx, dy, dz};
  double tmp = 0;
  for (int c = 0; c < ncomp; ++c)
    tmp += __longlong_as_double((long long)maxima[c]) / d[c];
  const double dtc = tmp > 0 ? cfl / tmp : dt_max;
  const double nw = fmin(dtc, dt_max);
  if (fabs(*dt - nw) / nw > 0.02) *dt = nw;
}

static __global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Integrating-factor linear combination over all modes of nvar fields,
// the building block of the Euler / RK2 / phase-shifted schemes
// (fluidsim/base/time_stepping/pseudo_spect.py:47-97, 245-796):
//   out = sum_j (a_j + b_j dt) exp(-e_j dt sigma(k)) exp(i p_j phi(k)) in_j
// with phi(k) = ax kx + ay ky + az kz (the phase shift of a translation by
// (ax, ay, az)), dt the device time step.  out may alias one input.
template <typename T>
struct IFCombine {
  const cpx<T>* in[4];
  double a[4], b[4], e[4];
  int p[4];
  int nin;
  cpx<T>* out;
  long long ke;  // elements per field
  int nvar;
  double ax, ay, az;
  Box box;
  Visc visc;
  const double* dt;
  unsigned int* nanflag;  // set when a value of field 0 is NaN, or null
};

template <typename T>
static __global__ __launch_bounds__(256) void k_if_combine(IFCombine<T> c) {
  const Box& bx = c.box;
  const double dt = *c.dt;
  const long long n = c.ke;
  bool nan = false;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int ikx = (int)(i % bx.nkx);
    const long long r = i / bx.nkx;
    const int iy = (int)(r % bx.ny), iz = (int)(r / bx.ny);
    const int iky = ksigned(iy, bx.ny);
    const int ikz = bx.dim == 3 ? ksigned(iz, bx.nz) : 0;
    const double sig = freq_diss(bx, c.visc, ikx, iky, ikz);
    const double phi = c.ax * ikx * bx.dkx + c.ay * iky * bx.dky + c.az * ikz * bx.dkz;
    const double cph = cos(phi), sph = sin(phi);
    double w[4];
    cpx<double> ph[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= c.nin) break;
      w[j] = (c.a[j] + c.b[j] * dt) * (c.e[j] != 0.0 ? exp(-c.e[j] * dt * sig) : 1.0);
      ph[j] = c.p[j] == 0 ? mk<double>(1.0, 0.0)
                          : mk<double>(cph, c.p[j] > 0 ? sph : -sph);
    }
    for (int v = 0; v < c.nvar; ++v) {
      double sx = 0.0, sy = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j >= c.nin) break;
        const cpx<T> x = c.in[j][v * n + i];
        const double xr = x.x, xi = x.y;
        sx += w[j] * (ph[j].x * xr - ph[j].y * xi);
        sy += w[j] * (ph[j].x * xi + ph[j].y * xr);
      }
      c.out[v * n + i] = mk<T>(T(sx), T(sy));
      if (v == 0 && (isnan(sx) || isnan(sy))) nan = true;
    }
  }
  if (nan && c.nanflag) atomicOr(c.nanflag, 1u);
}

// Resolution change of nf K fields (fill_field_fft_2d,
// fluidsim/base/init_fields.py:300-314; fill_field_fft_3d,
// fluidsim/util/mini_oper_modif_resol.py:26-53): output mode (o0, o1, ox)
// takes the input mode of the same signed wavenumbers when
// -m/2 < k <= m/2 on each signed axis (m = min of the two sizes; index
// <= m/2 is a non-negative wavenumber) and ox < min(nkx); the other output
// modes are left unchanged.  nz = 1 for 2D.
__device__ __forceinline__ int resize_src(int o, int n_out, int n_in) {
  const int m = n_out < n_in ? n_out : n_in;
  if (o <= m / 2) return o;
  const int s = o - n_out;  // negative wavenumber index
  return -s < m / 2 ? n_in + s : -1;
}

template <typename T>
static __global__ __launch_bounds__(256) void k_resize(const cpx<T>* in, int nz_in, int ny_in,
                                                int nkx_in, cpx<T>* out, int nz_out,
                                                int ny_out, int nkx_out, int nf) {
  const long long ko = (long long)nz_out * ny_out * nkx_out;
  const long long ki = (long long)nz_in * ny_in * nkx_in;
  const int mkx = nkx_out < nkx_in ? nkx_out : nkx_in;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ko * nf;
       i += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(i / ko);
    const long long j = i - f * ko;
    const int ox = (int)(j % nkx_out);
    const long long r = j / nkx_out;
    const int oy = (int)(r % ny_out), oz = (int)(r / ny_out);
    if (ox >= mkx) continue;
    const int sy = resize_src(oy, ny_out, ny_in);
    const int sz = nz_out == 1 && nz_in == 1 ? 0 : resize_src(oz, nz_out, nz_in);
    if (sy < 0 || sz < 0) continue;
    out[i] = in[f * ki + ((long long)sz * ny_in + sy) * nkx_in + ox];
  }
}

}  // namespace fsx
