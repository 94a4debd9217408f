"""Initial fields of the ns3d solver (fluidsim/solvers/ns3d/init_fields.py)."""

import numpy as np

from ...base.init_fields import InitFieldsBase


def compute_solenoidal_noise_fft(oper, length=None, velo_max=1.0, seed=42):
    """init_fields.py:153-196: legacy-seeded uniform noise in physical space,
    projected, dealiased, tanh band filter, rescaled to velo_max (engine
    transforms)."""
    lambda0 = oper.Lx / 4.0 if length is None else length

    def H_smooth(x, delta):
        return (1.0 + np.tanh(2 * np.pi * x / delta)) / 2.0

    np.random.seed(seed)
    vv = [np.random.random(oper.shapeX) - 0.5 for _ in range(3)]
    vv_fft = []
    for vi in vv:
        f = oper.fft(vi)
        f[0, 0, 0] = 0.0
        vv_fft.append(f)
    oper.project_perpk3d(*vv_fft)
    oper.dealiasing(*vv_fft)
    k0 = 2 * np.pi / lambda0
    K = np.sqrt(oper.K2)
    filt = H_smooth(k0 - K, 1.0 * k0)
    vv = [oper.ifft(vi * filt) for vi in vv_fft]
    vmax = np.sqrt(vv[0] ** 2 + vv[1] ** 2 + vv[2] ** 2).max()
    vv = [velo_max * vi / vmax for vi in vv]
    return tuple(oper.fft(vi) for vi in vv)


class InitFieldsNS3D(InitFieldsBase):
    available_types = ("in_script", "constant", "from_simul", "from_file", "noise")

    @classmethod
    def _complete_params_specific(cls, params):
        params.init_fields._set_child(
            "noise", attribs={"velo_max": 1.0, "length": None}
        )

    def __call__(self):
        self.sim._init_projection()
        super().__call__()

    def init_noise(self):
        p = self.params.init_fields.noise
        vv = compute_solenoidal_noise_fft(self.sim.oper, p.length, p.velo_max)
        self.sim.state.init_from_vxvyvzfft(*vv)
