// Forcing on a coarse spectral box (fluidsim/base/forcing/specific.py).
//
// fluidsim computes its random forcings on a small coarse grid, normalises
// them once per time step so that the injection rate is forcing_rate, and
// copies the coarse array into a full-size forcing array that is added to
// the tendencies at every Runge-Kutta stage.  Here the coarse array itself
// (Fc, nvar x coarse K grid) stays on the device and the Runge-Kutta
// kernels add Fc to the tendency of the modes inside the coarse box
// (rk_point), so no full-size forcing array is read per stage.
//
// k_forcing_norm runs once per step, after the CFL update and before the
// stage-0 algebra, on the state at the start of the step:
//   f_r = f1 - coef (f1 - f0)              (forcingc_from_f0f1, 852-868)
//   a = dt/2 sum_w |f_r|^2,  b = sum_w Re(conj(v) f_r)     (587-677)
//   R = root of a R^2 + b R - forcing_rate                  (679-726)
//   Fc = R f_r put into the forced key (vt_fft -> velocity through
//        vecfft_from_vtfft, fluidsim/operators/operators3d.py:978-992)
// where v is the forced variable of the state at the forced modes and
// sum_w the Parseval-weighted sum of the coarse grid.
#pragma once

#include "fsx_fft.h"

namespace fsx {

enum { FKEY_VAR = 0, FKEY_VT = 1 };

template <typename T>
struct ForcingArgs {
  const cpx<T>* S;       // state (nvar fields, full K grid)
  long long ke;          // elements per full K field
  cpx<T>* fc;            // coarse forcing box, nvar fields of fvs elements
  long long fvs;
  int nvar;
  const int* modes;      // coarse linear indices of the forced modes
  const double* w;       // Parseval weights of the forced modes
  const cpx<double>* f0; // raw random forcings of the key at the modes
  const cpx<double>* f1;
  int nmodes;
  int nyc, nzc, nkxc;    // coarse grid (nzc = 1 in 2D)
  int ny, nz, nkx;       // full grid
  double dkx, dky;
  int key_kind, key_var;
  double coef, rate;
  int which_root;        // 0 minabs, 1 first, 2 second, 3 positive
  const double* dt;
  // when set: coef, rate and which_root of the step in device memory
  // (par[0..2], written by k_forcing_par), so that a captured graph reads
  // the values of each replay
  const double* par;
};

// par[0..2] = coef, rate, which_root of the coming step
static __global__ void k_forcing_par(double* par, double coef, double rate, double which_root) {
  if (threadIdx.x == 0) {
    par[0] = coef;
    par[1] = rate;
    par[2] = which_root;
  }
}

template <typename T>
__device__ __forceinline__ void forced_mode(const ForcingArgs<T>& f, int m,
                                            int& c, long long& idx, int& sy,
                                            int& cx) {
  c = f.modes[m];
  cx = c % f.nkxc;
  const int r = c / f.nkxc;
  const int cy = r % f.nyc, cz = r / f.nyc;
  sy = cy <= f.nyc / 2 ? cy : cy - f.nyc;
  const int sz = f.nzc > 1 ? (cz <= f.nzc / 2 ? cz : cz - f.nzc) : 0;
  const int iy = sy >= 0 ? sy : sy + f.ny;
  const int iz = sz >= 0 ? sz : sz + f.nz;
  idx = ((long long)iz * f.ny + iy) * f.nkx + cx;
}

// sin(phi), cos(phi) of the horizontal wavevector, with kh^2 = 1e-14 at
// kh = 0 (fluidsim/operators/operators3d.py:961-992)
__device__ __forceinline__ void horiz_angles(double kx, double ky, double& sp,
                                             double& cp) {
  double kh2 = kx * kx + ky * ky;
  if (kh2 == 0.0) kh2 = 1e-14;
  const double tmp = sqrt(1.0 / kh2);
  cp = kx * tmp;
  sp = ky * tmp;
}

__device__ __forceinline__ cpx<double> mulc(cpx<double> a, cpx<double> b) {
  return mk<double>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Root of a R^2 + b R + c = 0 chosen like coef_normalization_from_abc:
// "first" = (-b + sqrt D) / 2a, "second" = (-b - sqrt D) / 2a, computed
// without cancellation; a negative D gives the complex pair (positive
// imaginary part first).
__device__ inline cpx<double> norm_root(double a, double b, double c, int which) {
  if (a == 0.0) return mk<double>(0.0, 0.0);
  const double d = b * b - 4.0 * a * c;
  cpx<double> r1, r2;
  if (d >= 0.0) {
    const double sq = sqrt(d);
    if (b >= 0.0) {
      const double q = -0.5 * (b + sq);
      r2 = mk<double>(q / a, 0.0);
      r1 = mk<double>(q != 0.0 ? c / q : 0.0, 0.0);
    } else {
      const double q = -0.5 * (b - sq);
      r1 = mk<double>(q / a, 0.0);
      r2 = mk<double>(c / q, 0.0);
    }
  } else {
    const double sq = sqrt(-d);
    r1 = mk<double>(-b / (2.0 * a), sq / (2.0 * a));
    r2 = mk<double>(r1.x, -r1.y);
  }
  const double m1 = r1.x * r1.x + r1.y * r1.y, m2 = r2.x * r2.x + r2.y * r2.y;
  switch (which) {
    case 1: return r1;
    case 2: return r2;
    case 3: return r2.x > 0.0 ? r2 : r1;
    default: return m2 < m1 ? r2 : r1;
  }
}

// One workgroup of 256 threads (the coarse box has a few hundred modes).
template <typename T>
static __global__ __launch_bounds__(256) void k_forcing_norm(ForcingArgs<T> f) {
  __shared__ double red_a[256], red_b[256];
  __shared__ double s_R[2];
  const double coef = f.par ? f.par[0] : f.coef;
  double pa = 0.0, pb = 0.0;
  for (int m = threadIdx.x; m < f.nmodes; m += blockDim.x) {
    int c, sy, cx;
    long long idx;
    forced_mode(f, m, c, idx, sy, cx);
    const cpx<double> f0 = f.f0[m], f1 = f.f1[m];
    const cpx<double> fr = mk<double>(f1.x - coef * (f1.x - f0.x),
                                      f1.y - coef * (f1.y - f0.y));
    cpx<double> v;
    if (f.key_kind == FKEY_VT) {
      const cpx<T> vx = f.S[idx], vy = f.S[f.ke + idx];
      double sp, cp;
      horiz_angles(cx * f.dkx, sy * f.dky, sp, cp);
      // vt = -i sin(phi) vx + i cos(phi) vy
      v = mk<double>(sp * vx.y - cp * vy.y, -sp * vx.x + cp * vy.x);
    } else {
      const cpx<T> a = f.S[f.key_var * f.ke + idx];
      v = mk<double>(a.x, a.y);
    }
    pa += f.w[m] * (fr.x * fr.x + fr.y * fr.y);
    pb += f.w[m] * (v.x * fr.x + v.y * fr.y);
  }
  red_a[threadIdx.x] = pa;
  red_b[threadIdx.x] = pb;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red_a[threadIdx.x] += red_a[threadIdx.x + s];
      red_b[threadIdx.x] += red_b[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double a = *f.dt / 2.0 * red_a[0];
    const double rate = f.par ? f.par[1] : f.rate;
    const int which = f.par ? (int)f.par[2] : f.which_root;
    const cpx<double> R = norm_root(a, red_b[0], -rate, which);
    s_R[0] = R.x;
    s_R[1] = R.y;
  }
  __syncthreads();
  const cpx<double> R = mk<double>(s_R[0], s_R[1]);
  for (int m = threadIdx.x; m < f.nmodes; m += blockDim.x) {
    int c, sy, cx;
    long long idx;
    forced_mode(f, m, c, idx, sy, cx);
    const cpx<double> f0 = f.f0[m], f1 = f.f1[m];
    const cpx<double> fr = mk<double>(f1.x - coef * (f1.x - f0.x),
                                      f1.y - coef * (f1.y - f0.y));
    const cpx<double> g = mulc(R, fr);
    if (f.key_kind == FKEY_VT) {
      double sp, cp;
      horiz_angles(cx * f.dkx, sy * f.dky, sp, cp);
      // vx = i sin(phi) vt, vy = -i cos(phi) vt, vz = 0
      f.fc[c] = mk<T>(T(-sp * g.y), T(sp * g.x));
      f.fc[f.fvs + c] = mk<T>(T(cp * g.y), T(-cp * g.x));
      f.fc[2 * f.fvs + c] = zero<T>();
    } else {
      f.fc[f.key_var * f.fvs + c] = mk<T>(T(g.x), T(g.y));
    }
  }
}

// out[v][full(c)] += scale * Fc[v][c] over the whole coarse box.
template <typename T>
static __global__ __launch_bounds__(256) void k_forcing_add(cpx<T>* out, long long ke,
                                                     const cpx<T>* fc,
                                                     long long fvs, int nvar,
                                                     int nyc, int nzc, int nkxc,
                                                     int ny, int nz, int nkx,
                                                     T scl) {
  const long long n = fvs * nvar;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int v = (int)(i / fvs);
    const int c = (int)(i - v * fvs);
    const int cx = c % nkxc, r = c / nkxc;
    const int cy = r % nyc, cz = r / nyc;
    const int sy = cy <= nyc / 2 ? cy : cy - nyc;
    const int sz = nzc > 1 ? (cz <= nzc / 2 ? cz : cz - nzc) : 0;
    const int iy = sy >= 0 ? sy : sy + ny;
    const int iz = sz >= 0 ? sz : sz + nz;
    const long long idx = ((long long)iz * ny + iy) * nkx + cx;
    const cpx<T> a = fc[i];
    out[v * ke + idx] = out[v * ke + idx] + mk<T>(scl * a.x, scl * a.y);
  }
}

}  // namespace fsx
