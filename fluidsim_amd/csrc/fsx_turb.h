// Smagorinsky turbulence model kernels
// (fluidsim/base/turb_model/smagorinsky.py:18-110,
//  fluidsim/base/turb_model/stress_tensor.py:15-44).
//
// One model evaluation on the velocity v (3 K fields) is
//   k_strain_rate : S_ij^ = i (k_i v_j + k_j v_i) / 2        (6 K fields)
//   c2r of the 6 fields (fsx_ifft)
//   k_smag_phys   : g_ij = 2 C_nuT |S| S_ij in place, |S| = sqrt(S_ij S_ij)
//   r2c of the 6 fields (fsx_fft)
//   k_smag_div    : f_i = 2i sum_j k_j g_ij^, added to the tendency
// with the six fields ordered (xx, yy, zz, yx, zx, zy) as in the reference.
// All three are pointwise, HBM-bound grid-stride kernels.
#pragma once

#include "fsx_fft.h"

namespace fsx {

struct KGrid {
  int nz, ny, nkx;
  double dkx, dky, dkz;
  __device__ __forceinline__ void k(long long i, double& kx, double& ky,
                                    double& kz) const {
    const int ikx = (int)(i % nkx);
    const long long r = i / nkx;
    const int iy = (int)(r % ny), iz = (int)(r / ny);
    kx = ikx * dkx;
    ky = ksigned(iy, ny) * dky;
    kz = ksigned(iz, nz) * dkz;
  }
};

template <typename T>
__device__ __forceinline__ cpx<T> i_times(cpx<T> a, double s) {
  // i s a
  return mk<T>(T(-s * a.y), T(s * a.x));
}

template <typename T>
static __global__ __launch_bounds__(256) void k_strain_rate(const cpx<T>* v, cpx<T>* S,
                                                     long long ke, KGrid g) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ke;
       i += (long long)gridDim.x * blockDim.x) {
    double kx, ky, kz;
    g.k(i, kx, ky, kz);
    const cpx<T> vx = v[i], vy = v[ke + i], vz = v[2 * ke + i];
    S[i] = i_times(vx, kx);
    S[ke + i] = i_times(vy, ky);
    S[2 * ke + i] = i_times(vz, kz);
    S[3 * ke + i] = i_times(vx, 0.5 * ky) + i_times(vy, 0.5 * kx);
    S[4 * ke + i] = i_times(vx, 0.5 * kz) + i_times(vz, 0.5 * kx);
    S[5 * ke + i] = i_times(vy, 0.5 * kz) + i_times(vz, 0.5 * ky);
  }
}

template <typename T>
static __global__ __launch_bounds__(256) void k_smag_phys(T* S, long long xe, T coef) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < xe;
       i += (long long)gridDim.x * blockDim.x) {
    const T sxx = S[i], syy = S[xe + i], szz = S[2 * xe + i];
    const T syx = S[3 * xe + i], szx = S[4 * xe + i], szy = S[5 * xe + i];
    const T nrm = sqrt(sxx * sxx + syy * syy + szz * szz +
                       T(2) * (syx * syx + szx * szx + szy * szy));
    const T c = coef * nrm;
    S[i] = c * sxx;
    S[xe + i] = c * syy;
    S[2 * xe + i] = c * szz;
    S[3 * xe + i] = c * syx;
    S[4 * xe + i] = c * szx;
    S[5 * xe + i] = c * szy;
  }
}

template <typename T>
static __global__ __launch_bounds__(256) void k_smag_div(const cpx<T>* gk, cpx<T>* out,
                                                  long long ke, KGrid g,
                                                  int accumulate) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ke;
       i += (long long)gridDim.x * blockDim.x) {
    double kx, ky, kz;
    g.k(i, kx, ky, kz);
    const cpx<T> gxx = gk[i], gyy = gk[ke + i], gzz = gk[2 * ke + i];
    const cpx<T> gyx = gk[3 * ke + i], gzx = gk[4 * ke + i], gzy = gk[5 * ke + i];
    // f = 2i (k . g), g symmetric
    const cpx<T> fx = i_times(gxx, 2 * kx) + i_times(gyx, 2 * ky) + i_times(gzx, 2 * kz);
    const cpx<T> fy = i_times(gyx, 2 * kx) + i_times(gyy, 2 * ky) + i_times(gzy, 2 * kz);
    const cpx<T> fz = i_times(gzx, 2 * kx) + i_times(gzy, 2 * ky) + i_times(gzz, 2 * kz);
    if (accumulate) {
      out[i] = out[i] + fx;
      out[ke + i] = out[ke + i] + fy;
      out[2 * ke + i] = out[2 * ke + i] + fz;
    } else {
      out[i] = fx;
      out[ke + i] = fy;
      out[2 * ke + i] = fz;
    }
  }
}

}  // namespace fsx
