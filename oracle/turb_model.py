"""FP64 restatement of fluidsim's Smagorinsky turbulence model.

Test infrastructure only (see ``oracle/__init__.py``).

Reference: fluidsim/base/turb_model/smagorinsky.py:18-110 (SmagorinskyModel)
and fluidsim/base/turb_model/stress_tensor.py:15-44
(StressTensorComputer3D); hooked into ns3d's tendencies at
fluidsim/solvers/ns3d/solver.py:246-249 (added before the projection).

  S_ij = (d_i v_j + d_j v_i) / 2,  |S| = sqrt(S_ij S_ij),
  g_ij = 2 C_nuT |S| S_ij,  C_nuT = C Delta^2 sqrt(2),  Delta = Lx / nx,
  f_i  = 2i sum_j k_j fft(g_ij).
"""

from math import sqrt

import numpy as np

from .operators import Oper3D


class Smagorinsky:
    def __init__(self, oper: Oper3D, C=0.18):
        self.oper = oper
        delta = oper.Lx / oper.nx
        self.C_nu_T = C * delta**2 * sqrt(2)

    def stress_tensor(self, vx_fft, vy_fft, vz_fft):
        """compute_stress_tensor (stress_tensor.py:27-39)."""
        o = self.oper
        ifft = o.ifft

        def grad(a):
            return (ifft(1j * o.Kx * a), ifft(1j * o.Ky * a), ifft(1j * o.Kz * a))

        dx_ux, dy_ux, dz_ux = grad(vx_fft)
        dx_uy, dy_uy, dz_uy = grad(vy_fft)
        dx_uz, dy_uz, dz_uz = grad(vz_fft)
        return (dx_ux, dy_uy, dz_uz, 0.5 * (dy_ux + dx_uy),
                0.5 * (dz_ux + dx_uz), 0.5 * (dz_uy + dy_uz))

    @staticmethod
    def norm(Sxx, Syy, Szz, Syx, Szx, Szy):
        """compute_norm (stress_tensor.py:41-44)."""
        return np.sqrt(Sxx**2 + Syy**2 + Szz**2 + 2 * (Syx**2 + Szx**2 + Szy**2))

    def get_forcing(self, vx_fft, vy_fft, vz_fft):
        """smagorinsky.py:52-110."""
        o = self.oper
        S = self.stress_tensor(vx_fft, vy_fft, vz_fft)
        nu_T_2 = 2 * self.C_nu_T * self.norm(*S)
        gxx, gyy, gzz, gyx, gzx, gzy = (o.fft(nu_T_2 * s) for s in S)
        Kx, Ky, Kz = o.Kx, o.Ky, o.Kz
        fx = 2j * (Kx * gxx + Ky * gyx + Kz * gzx)
        fy = 2j * (Kx * gyx + Ky * gyy + Kz * gzy)
        fz = 2j * (Kx * gzx + Ky * gzy + Kz * gzz)
        return fx, fy, fz
