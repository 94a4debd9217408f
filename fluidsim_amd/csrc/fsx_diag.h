// On-device diagnostics: spectra, spectral energy budgets and spatial
// means in one pass over the K arrays, accumulated in FP64.
//
// Per-mode quantities (one "kind" per solver output):
//   NS2D_SPECTRA  fluidsim/solvers/ns2d/output/spectra.py:16-28
//   NS2D_BUDGET   fluidsim/solvers/ns2d/output/spect_energy_budget.py:23-80
//   NS3D_SPECTRA  fluidsim/solvers/ns3d/output/spectra.py:15-99
//   STRAT_SPECTRA fluidsim/solvers/ns3d/strat/output/spectra.py:18-112
//   NS3D_BUDGET   fluidsim/solvers/ns3d/output/spect_energy_budget.py:106-190
//   STRAT_BUDGET  fluidsim/solvers/ns3d/strat/output/spect_energy_budget.py:60-118
//   NS2D_MEANS    fluidsim/solvers/ns2d/output/spatial_means.py:23-68
//   NS3D_MEANS    fluidsim/solvers/ns3d/output/spatial_means.py:23-77
//   STRAT_MEANS   fluidsim/solvers/ns3d/strat/output/spatial_means.py:27-106
// Each quantity q(k) is weighted by the Parseval weight of the half
// spectrum (1 on the kx = 0 plane and on the kx = nx/2 plane for even nx,
// 2 elsewhere) and binned into the histograms the flags ask for:
//   kx / ky / kz : bin |i_k| on that axis (negative ky, kz fold onto |k|),
//                  divided by dk of the axis (compute_1dspectra,
//                  spectra1D_from_fft);
//   shell        : |k| with dk = max(dk_i), floor bin j and linear share
//                  c = (|k| - j dk) / dk between bins j and j + 1, every mode
//                  at j >= nk - 1 in the last bin, divided by dk
//                  (compute_3dspectrum; in 2D spectrum2D_from_fft with
//                  dk = dkx, nk = nx/2 + 1);
//   kzkh         : bin |i_kz| times the shared bin of kh = sqrt(kx^2 + ky^2)
//                  (dkh = max(dkx, dky)), divided by dkz dkh
//                  (compute_spectrum_kzkh);
//   sums         : the Parseval sum of each quantity (sum_wavenumbers).
//
// One block owns one kz plane and a run of ky rows; its histograms live in
// LDS (FP64 atomics) and go to a per-block partial record, which
// k_diag_reduce sums over blocks in a fixed order (deterministic across
// blocks).  The kz and kzkh histograms use the fact that a block sees a
// single kz: the reducer gathers the blocks of the two planes +-kz.
#pragma once

#include "fsx_pointwise.h"

namespace fsx {

enum DiagKind {
  DIAG_NS2D_SPECTRA = 0,
  DIAG_NS2D_BUDGET = 1,
  DIAG_NS3D_SPECTRA = 2,
  DIAG_STRAT_SPECTRA = 3,
  DIAG_NS3D_BUDGET = 4,
  DIAG_STRAT_BUDGET = 5,
  DIAG_NS2D_MEANS = 6,
  DIAG_NS3D_MEANS = 7,
  DIAG_STRAT_MEANS = 8,
  DIAG_NKINDS = 9
};

enum DiagFlag {
  DIAG_KX = 1,
  DIAG_KY = 2,
  DIAG_KZ = 4,
  DIAG_SHELL = 8,
  DIAG_KZKH = 16
};

__host__ __device__ constexpr int diag_nq(int kind) {
  return kind == DIAG_NS2D_SPECTRA    ? 1
         : kind == DIAG_NS2D_BUDGET   ? 2
         : kind == DIAG_NS3D_SPECTRA  ? 5
         : kind == DIAG_STRAT_SPECTRA ? 6
         : kind == DIAG_NS3D_BUDGET   ? 5
         : kind == DIAG_STRAT_BUDGET  ? 8
         : kind == DIAG_NS2D_MEANS    ? 10
         : kind == DIAG_NS3D_MEANS    ? 9
         : kind == DIAG_STRAT_MEANS   ? 18
                                      : 0;
}

// number of complex input fields each kind reads (forcing fields of the
// means kinds are optional: null pointers give zero injection terms)
__host__ __device__ constexpr int diag_nin(int kind) {
  return kind == DIAG_NS2D_SPECTRA    ? 1
         : kind == DIAG_NS2D_BUDGET   ? 4
         : kind == DIAG_NS3D_SPECTRA  ? 3
         : kind == DIAG_STRAT_SPECTRA ? 4
         : kind == DIAG_NS3D_BUDGET   ? 6
         : kind == DIAG_STRAT_BUDGET  ? 8
         : kind == DIAG_NS2D_MEANS    ? 2
         : kind == DIAG_NS3D_MEANS    ? 6
         : kind == DIAG_STRAT_MEANS   ? 8
                                      : 0;
}

struct DiagGeom {
  int dim, nx, ny, nz, nkx;
  double dkx, dky, dkz;
  int nkxs, nkys, nkzs;  // 1D bins: nx/2 + 1, ny/2 + 1, nz/2 + 1
  int nk3, nkh;          // shell and kh bins
  double dk3, dkh;
  int flags;
  int nchunk, rows;  // blocks = nz * nchunk, rows of ky per block
  int q0, nq;        // quantity group handled by this launch
  // offsets (doubles) in a block record, and its size
  int off_kx, off_ky, off_k3, off_kh, off_sum, rec;
  // the shell and kh histograms ([off_k3, off_sum) of the record) have
  // ncopy copies in LDS: lane l adds into copy l mod ncopy (neighbouring
  // lanes, often in one bin, then hit different addresses); copies 1.. sit
  // after the record and are summed into it before it is written
  int ncopy;
};

struct DiagIn {
  const void* f[8];
  double inv_N2;
  Visc visc;
};

struct Mode {
  double kx, ky, kz, k2;
  int ikx, iky, ikz;
};

template <typename T>
__device__ __forceinline__ cpx<double> ld(const void* p, size_t i) {
  const cpx<T> a = reinterpret_cast<const cpx<T>*>(p)[i];
  return cpx<double>{(double)a.x, (double)a.y};
}

__device__ __forceinline__ double abs2(cpx<double> a) { return a.x * a.x + a.y * a.y; }
// Re(conj(a) b)
__device__ __forceinline__ double re_cj(cpx<double> a, cpx<double> b) {
  return a.x * b.x + a.y * b.y;
}
__device__ __forceinline__ cpx<double> cscale(cpx<double> a, double s) {
  return cpx<double>{a.x * s, a.y * s};
}
__device__ __forceinline__ cpx<double> csub(cpx<double> a, cpx<double> b) {
  return cpx<double>{a.x - b.x, a.y - b.y};
}
// i s a
__device__ __forceinline__ cpx<double> ci(cpx<double> a, double s) {
  return cpx<double>{-s * a.y, s * a.x};
}

// ux = i ky rot / K2_not0, uy = -i kx rot / K2_not0 with K2_not0(0) = 1e-14
// (fluidfft vecfft_from_rotfft)
__device__ __forceinline__ void vel2d(cpx<double> rot, const Mode& m, cpx<double>& ux,
                                      cpx<double>& uy) {
  const double inv = 1.0 / (m.k2 == 0 ? 1e-14 : m.k2);
  ux = ci(rot, m.ky * inv);
  uy = ci(rot, -m.kx * inv);
}

// urudfft_from_vxvyfft (fluidsim/operators/operators3d.py:713-730)
__device__ __forceinline__ void urud(cpx<double> vx, cpx<double> vy, const Mode& m,
                                     cpx<double>& urx, cpx<double>& ury,
                                     cpx<double>& udx, cpx<double>& udy) {
  double kh2 = m.kx * m.kx + m.ky * m.ky;
  if (kh2 == 0) kh2 = 1e-14;
  const double inv = 1.0 / kh2;
  const cpx<double> kdotu = cpx<double>{m.kx * vx.x + m.ky * vy.x,
                                        m.kx * vx.y + m.ky * vy.y};
  udx = cscale(kdotu, m.kx * inv);
  udy = cscale(kdotu, m.ky * inv);
  urx = csub(vx, udx);
  ury = csub(vy, udy);
}

__device__ __forceinline__ void freqs(const DiagIn& in, const Box& b, const Mode& m,
                                      double& fd, double& fh) {
  Visc vs = in.visc;
  vs.num4 = 0.0;
  fd = freq_diss(b, vs, m.ikx, m.iky, m.ikz);
  if (in.visc.num4 != 0) {
    const Visc vh = {0.0, 0.0, 0.0, in.visc.num4};
    fh = freq_diss(b, vh, m.ikx, m.iky, m.ikz);
  } else {
    fh = 0.0;
  }
}

// Per-mode quantities of a kind (unweighted).
template <int KIND, typename T>
__device__ __forceinline__ void diag_eval(const DiagIn& in, const Box& b, size_t i,
                                          const Mode& m, double* q) {
  if constexpr (KIND == DIAG_NS2D_SPECTRA) {
    cpx<double> ux, uy;
    vel2d(ld<T>(in.f[0], i), m, ux, uy);
    q[0] = 0.5 * (abs2(ux) + abs2(uy));
  } else if constexpr (KIND == DIAG_NS2D_BUDGET) {
    const cpx<double> rot = ld<T>(in.f[0], i);
    cpx<double> ux, uy;
    vel2d(rot, m, ux, uy);
    q[0] = re_cj(ux, ld<T>(in.f[2], i)) + re_cj(uy, ld<T>(in.f[3], i));
    q[1] = re_cj(rot, ld<T>(in.f[1], i));
  } else if constexpr (KIND == DIAG_NS3D_SPECTRA || KIND == DIAG_STRAT_SPECTRA) {
    const cpx<double> vx = ld<T>(in.f[0], i), vy = ld<T>(in.f[1], i),
                      vz = ld<T>(in.f[2], i);
    cpx<double> urx, ury, udx, udy;
    urud(vx, vy, m, urx, ury, udx, udy);
    q[0] = 0.5 * abs2(vx);
    q[1] = 0.5 * abs2(vy);
    q[2] = 0.5 * abs2(vz);
    int o = 3;
    if constexpr (KIND == DIAG_STRAT_SPECTRA) {
      q[3] = 0.5 * in.inv_N2 * abs2(ld<T>(in.f[3], i));
      o = 4;
    }
    q[o] = 0.5 * (abs2(urx) + abs2(ury));
    q[o + 1] = 0.5 * (abs2(udx) + abs2(udy));
  } else if constexpr (KIND == DIAG_NS3D_BUDGET || KIND == DIAG_STRAT_BUDGET) {
    const cpx<double> vx = ld<T>(in.f[0], i), vy = ld<T>(in.f[1], i),
                      vz = ld<T>(in.f[2], i);
    const cpx<double> fx = ld<T>(in.f[3], i), fy = ld<T>(in.f[4], i),
                      fz = ld<T>(in.f[5], i);
    cpx<double> urx, ury, udx, udy;
    urud(vx, vy, m, urx, ury, udx, udy);
    double fd, fh;
    freqs(in, b, m, fd, fh);
    q[0] = re_cj(vx, fx) + re_cj(vy, fy);    // transfer_Kh
    q[1] = re_cj(vz, fz);                    // transfer_Kz
    q[2] = re_cj(urx, fx) + re_cj(ury, fy);  // transfer_Khr
    q[3] = fd * (abs2(vx) + abs2(vy));       // diss_Kh
    q[4] = fd * abs2(vz);                    // diss_Kz
    if constexpr (KIND == DIAG_STRAT_BUDGET) {
      const cpx<double> bb = ld<T>(in.f[6], i);
      // f[7] is the b tendency -i k.(v b)^ - N^2 vz; the budget's
      // fb = -i k.(v b)^ / N^2 = f[7] / N^2 + vz
      const cpx<double> ft = ld<T>(in.f[7], i);
      const cpx<double> fb = cpx<double>{ft.x * in.inv_N2 + vz.x, ft.y * in.inv_N2 + vz.y};
      q[5] = in.inv_N2 * fd * abs2(bb);  // diss_A
      q[6] = -re_cj(vz, bb);             // conv_K2A
      q[7] = re_cj(bb, fb);              // transfer_A
    }
  } else if constexpr (KIND == DIAG_NS2D_MEANS) {
    const cpx<double> rot = ld<T>(in.f[0], i);
    cpx<double> ux, uy;
    vel2d(rot, m, ux, uy);
    double fd, fh;
    freqs(in, b, m, fd, fh);
    const double e = 0.5 * (abs2(ux) + abs2(uy)), z = 0.5 * abs2(rot);
    q[0] = e;
    q[1] = z;
    q[2] = fd * 2 * e;
    q[3] = fh * 2 * e;
    q[4] = fd * 2 * z;
    q[5] = fh * 2 * z;
    if (in.f[1]) {
      const cpx<double> fr = ld<T>(in.f[1], i);
      cpx<double> fx, fy;
      vel2d(fr, m, fx, fy);
      q[6] = re_cj(ux, fx) + re_cj(uy, fy);  // PK1
      q[7] = abs2(fx) + abs2(fy);            // PK2 / (dt / 2)
      q[8] = re_cj(rot, fr);                 // PZ1
      q[9] = abs2(fr);                       // PZ2 / (dt / 2)
    } else {
      q[6] = q[7] = q[8] = q[9] = 0.0;
    }
  } else if constexpr (KIND == DIAG_NS3D_MEANS) {
    const cpx<double> vx = ld<T>(in.f[0], i), vy = ld<T>(in.f[1], i),
                      vz = ld<T>(in.f[2], i);
    double fd, fh;
    freqs(in, b, m, fd, fh);
    q[0] = 0.5 * abs2(vx);
    q[1] = 0.5 * abs2(vy);
    q[2] = 0.5 * abs2(vz);
    const double e2 = 2 * (q[0] + q[1] + q[2]);
    const double k4 = m.k2 * m.k2;
    q[3] = fd * e2;
    q[4] = fh * e2;
    q[5] = in.visc.nu4 * k4 * e2;
    q[6] = in.visc.nu8 * (k4 * k4) * e2;
    if (in.f[3]) {
      const cpx<double> fx = ld<T>(in.f[3], i), fy = ld<T>(in.f[4], i),
                        fz = ld<T>(in.f[5], i);
      q[7] = re_cj(vx, fx) + re_cj(vy, fy) + re_cj(vz, fz);
      q[8] = abs2(fx) + abs2(fy) + abs2(fz);
    } else {
      q[7] = q[8] = 0.0;
    }
  } else if constexpr (KIND == DIAG_STRAT_MEANS) {
    const cpx<double> vx = ld<T>(in.f[0], i), vy = ld<T>(in.f[1], i),
                      vz = ld<T>(in.f[2], i), bb = ld<T>(in.f[3], i);
    cpx<double> urx, ury, udx, udy;
    urud(vx, vy, m, urx, ury, udx, udy);
    double fd, fh;
    freqs(in, b, m, fd, fh);
    const double A = 0.5 * in.inv_N2 * abs2(bb);
    const double Kz = 0.5 * abs2(vz);
    const double Khr = 0.5 * (abs2(urx) + abs2(ury));
    const double Khd = 0.5 * (abs2(udx) + abs2(udy));
    const bool shear = m.kx * m.kx + m.ky * m.ky == 0.0;
    const double eK2 = 2 * (Kz + Khr + Khd), eA2 = 2 * A;
    const double k4 = m.k2 * m.k2;
    q[0] = A;
    q[1] = Kz;
    q[2] = Khr;
    q[3] = Khd;
    q[4] = shear ? Khr : 0.0;  // Khs
    q[5] = shear ? A : 0.0;    // As
    q[6] = fd * eK2;
    q[7] = fh * eK2;
    q[8] = fd * eA2;
    q[9] = fh * eA2;
    q[10] = in.visc.nu4 * k4 * eK2;
    q[11] = in.visc.nu4 * k4 * eA2;
    q[12] = in.visc.nu8 * (k4 * k4) * eK2;
    q[13] = in.visc.nu8 * (k4 * k4) * eA2;
    if (in.f[4]) {
      const cpx<double> fx = ld<T>(in.f[4], i), fy = ld<T>(in.f[5], i),
                        fz = ld<T>(in.f[6], i), fb = ld<T>(in.f[7], i);
      q[14] = re_cj(vx, fx) + re_cj(vy, fy) + re_cj(vz, fz);  // PK1
      q[15] = abs2(fx) + abs2(fy) + abs2(fz);                 // PK2 / (dt / 2)
      q[16] = re_cj(bb, fb);                                  // PA1 N^2
      q[17] = abs2(fb);                                       // PA2 N^2 / (dt / 2)
    } else {
      q[14] = q[15] = q[16] = q[17] = 0.0;
    }
  }
}

// floor bin with linear share; every value at j >= n - 1 goes to n - 1
__device__ __forceinline__ void share_bin(double k, double dk, int n, int& j, double& c) {
  j = (int)(k / dk);
  if (j >= n - 1) {
    j = n - 1;
    c = 0.0;
  } else {
    c = (k - j * dk) / dk;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// NT threads per block (waves take the block's ky rows in turn): 256, or
// 512 for records that leave room for one block per CU
template <typename T, int KIND, int NT>
static __global__ __launch_bounds__(NT) void k_diag(DiagIn in, Box box, DiagGeom g,
                                              double* partial) {
#pragma clang fp contract(off)
  constexpr int Q = diag_nq(KIND);
  extern __shared__ double lds[];
  const int span = g.off_sum - g.off_k3;
  const int lds_n = g.rec + (g.ncopy > 1 ? (g.ncopy - 1) * span : 0);
  for (int t = threadIdx.x; t < lds_n; t += blockDim.x) lds[t] = 0.0;
  __syncthreads();
  const int blk = blockIdx.x;
  const int iz = blk / g.nchunk, ch = blk % g.nchunk;
  const int y0 = ch * g.rows;
  const int y1 = min(g.ny, y0 + g.rows);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  const bool even = (g.nx & 1) == 0;
  // this lane's copy of the shell / kh histograms: record offset + shift
  const int cp = g.ncopy > 1 ? lane % g.ncopy : 0;
  const int hshift = cp == 0 ? 0 : g.rec + (cp - 1) * span - g.off_k3;
  Mode m;
  m.ikz = g.dim == 3 ? ksigned(iz, g.nz) : 0;
  m.kz = m.ikz * g.dkz;
  double tot[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s) tot[s] = 0.0;
  // With shell or kh histograms, lanes own runs of kDiagRun consecutive kx
  // (blocks of 64 kDiagRun modes): along a run |k| and kh grow slowly, so
  // each lane sums the shares of its modes in registers while the bin stays
  // the same and adds them to the LDS histograms when it moves on (and at
  // the end of the row), instead of two LDS atomics per mode that
  // neighbouring lanes aim at the same bin.  A lane's loads are then
  // kDiagRun elements apart (the block is still read in full).  Without
  // those histograms lanes take consecutive kx.
  // (kinds of at most 6 quantities: the means kinds carry 9-18 and bin
  // nothing, the strat budget's 8 running sums would spill; those keep one
  // lane per kx and direct atomics)
  constexpr int kDiagRun = 4;
  constexpr bool kRuns = Q <= 6;
  const bool runs = kRuns && (g.flags & (DIAG_SHELL | DIAG_KZKH)) != 0;
  for (int iy = y0 + wid; iy < y1; iy += nw) {
    m.iky = ksigned(iy, g.ny);
    m.ky = m.iky * g.dky;
    double row[Q];
#pragma unroll
    for (int s = 0; s < Q; ++s) row[s] = 0.0;
    // running bins (-1: none yet) and the sums for bins j and j + 1
    constexpr int QR = kRuns ? Q : 1;
    int js = -1, jh = -1;
    double s0[QR], s1[QR], h0[QR], h1[QR];
#pragma unroll
    for (int s = 0; s < QR; ++s) s0[s] = s1[s] = h0[s] = h1[s] = 0.0;
    // adds (a0, a1) to bins (j, j + 1) of histogram off (n bins)
    auto flush = [&](int off, int n, int j, const double(&a0)[QR], const double(&a1)[QR]) {
#pragma unroll
      for (int s = 0; s < Q; ++s)
        if (s >= g.q0 && s < g.q0 + g.nq) {
          double* h = &lds[hshift + off + (s - g.q0) * n];
          atomicAdd(h + j, a0[s]);
          if (j + 1 < n && a1[s] != 0.0) atomicAdd(h + j + 1, a1[s]);
        }
    };
    // mode with bin j and share c of value q into the running sums
    auto accumulate = [&](int off, int n, int j, double c, const double(&q)[Q], int& jc,
                          double(&a0)[QR], double(&a1)[QR]) {
      if (j != jc) {
        if (jc >= 0 && j == jc + 1) {
          // bin jc is complete; its successor's sum so far moves down
#pragma unroll
          for (int s = 0; s < Q; ++s)
            if (s >= g.q0 && s < g.q0 + g.nq)
              atomicAdd(&lds[hshift + off + (s - g.q0) * n + jc], a0[s]);
#pragma unroll
          for (int s = 0; s < Q; ++s) {
            a0[s] = a1[s];
            a1[s] = 0.0;
          }
        } else {
          if (jc >= 0) flush(off, n, jc, a0, a1);
#pragma unroll
          for (int s = 0; s < Q; ++s) a0[s] = a1[s] = 0.0;
        }
        jc = j;
      }
#pragma unroll
      for (int s = 0; s < Q; ++s) {
        a0[s] += (1.0 - c) * q[s];
        a1[s] += c * q[s];
      }
    };
    auto mode = [&](int ikx) {
      Mode mm = m;  // (a copy: the row's m stays out of the lambda's reach)
      mm.ikx = ikx;
      mm.kx = ikx * g.dkx;
      mm.k2 = mm.kx * mm.kx + mm.ky * mm.ky + mm.kz * mm.kz;
      const size_t i = ((size_t)iz * g.ny + iy) * g.nkx + ikx;
      const double w = (ikx == 0 || (even && ikx == g.nx / 2)) ? 1.0 : 2.0;
      double q[Q];
      diag_eval<KIND, T>(in, box, i, mm, q);
#pragma unroll
      for (int s = 0; s < Q; ++s) {
        q[s] *= w;
        row[s] += q[s];
      }
      if (g.flags & DIAG_KX) {
#pragma unroll
        for (int s = 0; s < Q; ++s)
          if (s >= g.q0 && s < g.q0 + g.nq)
            atomicAdd(&lds[g.off_kx + (s - g.q0) * g.nkxs + ikx], q[s]);
      }
      if (g.flags & DIAG_SHELL) {
        int j;
        double c;
        share_bin(sqrt(mm.k2), g.dk3, g.nk3, j, c);
        accumulate(g.off_k3, g.nk3, j, c, q, js, s0, s1);
      }
      if (g.flags & DIAG_KZKH) {
        int j;
        double c;
        share_bin(sqrt(mm.kx * mm.kx + mm.ky * mm.ky), g.dkh, g.nkh, j, c);
        accumulate(g.off_kh, g.nkh, j, c, q, jh, h0, h1);
      }
    };
    if constexpr (kRuns) {
      if (runs) {
        for (int blk0 = 0; blk0 < g.nkx; blk0 += kDiagRun * 64)
#pragma unroll
          for (int r = 0; r < kDiagRun; ++r) {
            const int ikx = blk0 + lane * kDiagRun + r;
            if (ikx < g.nkx) mode(ikx);
          }
      } else {
        for (int ikx = lane; ikx < g.nkx; ikx += 64) mode(ikx);
      }
    } else {
      for (int ikx = lane; ikx < g.nkx; ikx += 64) {
        m.ikx = ikx;
        m.kx = ikx * g.dkx;
        m.k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
        const size_t i = ((size_t)iz * g.ny + iy) * g.nkx + ikx;
        const double w = (ikx == 0 || (even && ikx == g.nx / 2)) ? 1.0 : 2.0;
        double q[Q];
        diag_eval<KIND, T>(in, box, i, m, q);
#pragma unroll
        for (int s = 0; s < Q; ++s) {
          q[s] *= w;
          row[s] += q[s];
        }
        if (g.flags & DIAG_KX) {
#pragma unroll
          for (int s = 0; s < Q; ++s)
            if (s >= g.q0 && s < g.q0 + g.nq)
              atomicAdd(&lds[g.off_kx + (s - g.q0) * g.nkxs + ikx], q[s]);
        }
        if (g.flags & DIAG_SHELL) {
          int j;
          double c;
          share_bin(sqrt(m.k2), g.dk3, g.nk3, j, c);
#pragma unroll
          for (int s = 0; s < Q; ++s)
            if (s >= g.q0 && s < g.q0 + g.nq) {
              double* h = &lds[hshift + g.off_k3 + (s - g.q0) * g.nk3];
              atomicAdd(h + j, (1.0 - c) * q[s]);
              if (c != 0.0) atomicAdd(h + j + 1, c * q[s]);
            }
        }
        if (g.flags & DIAG_KZKH) {
          int j;
          double c;
          share_bin(sqrt(m.kx * m.kx + m.ky * m.ky), g.dkh, g.nkh, j, c);
#pragma unroll
          for (int s = 0; s < Q; ++s)
            if (s >= g.q0 && s < g.q0 + g.nq) {
              double* h = &lds[hshift + g.off_kh + (s - g.q0) * g.nkh];
              atomicAdd(h + j, (1.0 - c) * q[s]);
              if (c != 0.0) atomicAdd(h + j + 1, c * q[s]);
            }
        }
      }
    }
    if constexpr (kRuns) {
      if (js >= 0) flush(g.off_k3, g.nk3, js, s0, s1);
      if (jh >= 0) flush(g.off_kh, g.nkh, jh, h0, h1);
    }
    if (g.flags & DIAG_KY) {
      const int jy = m.iky < 0 ? -m.iky : m.iky;
#pragma unroll
      for (int s = 0; s < Q; ++s)
        if (s >= g.q0 && s < g.q0 + g.nq) {
          const double v = wave_sum(row[s]);
          if (lane == 0) atomicAdd(&lds[g.off_ky + (s - g.q0) * g.nkys + jy], v);
        }
    }
#pragma unroll
    for (int s = 0; s < Q; ++s) tot[s] += row[s];
  }
#pragma unroll
  for (int s = 0; s < Q; ++s)
    if (s >= g.q0 && s < g.q0 + g.nq) {
      const double v = wave_sum(tot[s]);
      if (lane == 0) atomicAdd(&lds[g.off_sum + (s - g.q0)], v);
    }
  __syncthreads();
  if (g.ncopy > 1) {
    for (int t = threadIdx.x; t < span; t += blockDim.x) {
      double v = lds[g.off_k3 + t];
      for (int c = 1; c < g.ncopy; ++c) v += lds[g.rec + (c - 1) * span + t];
      lds[g.off_k3 + t] = v;
    }
    __syncthreads();
  }
  double* dst = partial + (size_t)blk * g.rec;
  for (int t = threadIdx.x; t < g.rec; t += blockDim.x) dst[t] = lds[t];
}

// Sum of the block records into the output of one quantity group.
// Output layout for the whole kind (Q = all quantities):
//   [kx Q*nkxs][ky Q*nkys][kz Q*nkzs][shell Q*nk3][kzkh Q*nkzs*nkh][sums Q]
// (absent histograms take no room).
struct DiagOut {
  long long o_kx, o_ky, o_kz, o_k3, o_kzkh, o_sum;
  int Q;
};

// One wave per output value: the lanes sum strided subsets of the
// contributing block records, then a fixed shuffle tree adds the lanes
// (deterministic for a given grid).
static __global__ __launch_bounds__(256) void k_diag_reduce(const double* partial, DiagGeom g,
                                                     DiagOut o, double* out) {
  const int nq = g.nq;
  const long long n_kx = (g.flags & DIAG_KX) ? (long long)nq * g.nkxs : 0;
  const long long n_ky = (g.flags & DIAG_KY) ? (long long)nq * g.nkys : 0;
  const long long n_kz = (g.flags & DIAG_KZ) ? (long long)nq * g.nkzs : 0;
  const long long n_k3 = (g.flags & DIAG_SHELL) ? (long long)nq * g.nk3 : 0;
  const long long n_kh = (g.flags & DIAG_KZKH) ? (long long)nq * g.nkzs * g.nkh : 0;
  const long long n_sum = nq;
  const long long total = n_kx + n_ky + n_kz + n_k3 + n_kh + n_sum;
  const int nb = g.nz * g.nchunk;
  const int lane = threadIdx.x & 63;
  const long long nwaves = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long t = blockIdx.x * (long long)(blockDim.x >> 6) + (threadIdx.x >> 6);
       t < total; t += nwaves) {
    long long e = t;
    // contributing records: every block, or (jz >= 0) the blocks of the
    // planes iz = jz and iz = nz - jz
    int src = 0, jz = -1;
    long long dst = 0;
    double scale = 1.0;
    if (e < n_kx) {
      const int s = (int)(e / g.nkxs), j = (int)(e % g.nkxs);
      src = g.off_kx + (int)e;
      dst = o.o_kx + (long long)(g.q0 + s) * g.nkxs + j;
      scale = 1.0 / g.dkx;
    } else if ((e -= n_kx) < n_ky) {
      const int s = (int)(e / g.nkys), j = (int)(e % g.nkys);
      src = g.off_ky + (int)e;
      dst = o.o_ky + (long long)(g.q0 + s) * g.nkys + j;
      scale = 1.0 / g.dky;
    } else if ((e -= n_ky) < n_kz) {
      const int s = (int)(e / g.nkzs);
      jz = (int)(e % g.nkzs);
      src = g.off_sum + s;
      dst = o.o_kz + (long long)(g.q0 + s) * g.nkzs + jz;
      scale = 1.0 / g.dkz;
    } else if ((e -= n_kz) < n_k3) {
      const int s = (int)(e / g.nk3), j = (int)(e % g.nk3);
      src = g.off_k3 + (int)e;
      dst = o.o_k3 + (long long)(g.q0 + s) * g.nk3 + j;
      scale = 1.0 / g.dk3;
    } else if ((e -= n_k3) < n_kh) {
      const long long per = (long long)g.nkzs * g.nkh;
      const int s = (int)(e / per);
      jz = (int)((e % per) / g.nkh);
      const int jh = (int)(e % g.nkh);
      src = g.off_kh + s * g.nkh + jh;
      dst = o.o_kzkh + (long long)(g.q0 + s) * per + (long long)jz * g.nkh + jh;
      scale = 1.0 / (g.dkz * g.dkh);
    } else {
      e -= n_kh;
      src = g.off_sum + (int)e;
      dst = o.o_sum + g.q0 + e;
    }
    double acc = 0.0;
    if (jz < 0) {
      for (int b = lane; b < nb; b += 64) acc += partial[(size_t)b * g.rec + src];
    } else {
      const int iz2 = (jz == 0 || 2 * jz == g.nz) ? -1 : g.nz - jz;
      const int m = (iz2 >= 0 ? 2 : 1) * g.nchunk;
      for (int k = lane; k < m; k += 64) {
        const int b = k < g.nchunk ? jz * g.nchunk + k : iz2 * g.nchunk + (k - g.nchunk);
        acc += partial[(size_t)b * g.rec + src];
      }
    }
    acc = wave_sum(acc);
    if (lane == 0) out[dst] = acc * scale;
  }
}

// ns2d budget fields (spect_energy_budget.py:27-53): from rot_fft the eight
// spectral fields ux, uy, d_x rot, d_y rot, d_x ux, d_y ux, d_x uy, d_y uy
// (gradfft_from_fft: i kx a, i ky a) over all modes.
template <typename T>
static __global__ __launch_bounds__(256) void k_ns2d_budget_grads(const cpx<T>* rot, cpx<T>* w,
                                                           long long ke, int nkx, int ny,
                                                           double dkx, double dky) {
#pragma clang fp contract(off)
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ke;
       i += (long long)gridDim.x * blockDim.x) {
    Mode m;
    m.ikx = (int)(i % nkx);
    m.iky = ksigned((int)(i / nkx), ny);
    m.ikz = 0;
    m.kx = m.ikx * dkx;
    m.ky = m.iky * dky;
    m.kz = 0;
    m.k2 = m.kx * m.kx + m.ky * m.ky;
    const cpx<T> r0 = rot[i];
    const cpx<double> r = {(double)r0.x, (double)r0.y};
    cpx<double> ux, uy;
    vel2d(r, m, ux, uy);
    const cpx<double> v[8] = {ux,           uy,           ci(r, m.kx),  ci(r, m.ky),
                              ci(ux, m.kx), ci(ux, m.ky), ci(uy, m.kx), ci(uy, m.ky)};
#pragma unroll
    for (int f = 0; f < 8; ++f) w[f * ke + i] = mk<T>(T(v[f].x), T(v[f].y));
  }
}

// Frot = -ux d_x rot - uy (d_y rot + beta), Fx = -ux d_x ux - uy d_y ux,
// Fy = -ux d_x uy - uy d_y uy, written over fields 0..2 of x.
template <typename T>
static __global__ __launch_bounds__(256) void k_ns2d_budget_prod(T* x, long long xe, double beta) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < xe;
       i += (long long)gridDim.x * blockDim.x) {
    const T ux = x[i], uy = x[xe + i];
    const T prx = x[2 * xe + i], pry = x[3 * xe + i];
    const T pxux = x[4 * xe + i], pyux = x[5 * xe + i];
    const T pxuy = x[6 * xe + i], pyuy = x[7 * xe + i];
    x[i] = -ux * prx - uy * (pry + T(beta));
    x[xe + i] = -ux * pxux - uy * pyux;
    x[2 * xe + i] = -ux * pxuy - uy * pyuy;
  }
}

}  // namespace fsx
