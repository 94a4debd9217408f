"""Initial fields of the ns2d solver (fluidsim/solvers/ns2d/init_fields.py)."""

import numpy as np

from ...base.init_fields import InitFieldsBase


class InitFieldsNS2D(InitFieldsBase):
    available_types = ("in_script", "constant", "from_simul", "from_file", "noise")

    @classmethod
    def _complete_params_specific(cls, params):
        params.init_fields._set_child(
            "noise", attribs={"velo_max": 1.0, "length": 0.0}
        )

    def init_noise(self):
        """InitFieldsNoise.compute_rotuxuy_fft (init_fields.py:57-108):
        legacy-seeded uniform noise on the K grid, projected, dealiased,
        tanh band filter at k0 = 2 pi / lambda0, rescaled to velo_max.
        The transforms run on the engine."""
        params = self.params
        oper = self.sim.oper
        lambda0 = params.init_fields.noise.length
        if lambda0 == 0:
            lambda0 = min(oper.Lx, oper.Ly) / 4.0

        def H_smooth(x, delta):
            return (1.0 + np.tanh(2 * np.pi * x / delta)) / 2.0

        np.random.seed(42)
        ux_fft = (np.random.random(oper.shapeK) + 1j * np.random.random(oper.shapeK)
                  - 0.5 - 0.5j)
        uy_fft = (np.random.random(oper.shapeK) + 1j * np.random.random(oper.shapeK)
                  - 0.5 - 0.5j)
        ux_fft[0, 0] = 0.0
        uy_fft[0, 0] = 0.0
        oper.projection_perp(ux_fft, uy_fft)
        oper.dealiasing(ux_fft, uy_fft)
        k0 = 2 * np.pi / lambda0
        filt = H_smooth(k0 - oper.K, 1.0 * k0)
        ux = oper.ifft2(ux_fft * filt)
        uy = oper.ifft2(uy_fft * filt)
        velo_max = np.sqrt(ux**2 + uy**2).max()
        ux = params.init_fields.noise.velo_max * ux / velo_max
        uy = params.init_fields.noise.velo_max * uy / velo_max
        rot_fft = oper.rotfft_from_vecfft(oper.fft2(ux), oper.fft2(uy))
        self.sim.state.init_from_rotfft(rot_fft)
