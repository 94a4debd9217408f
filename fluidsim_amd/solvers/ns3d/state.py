"""State of the ns3d solver (fluidsim/solvers/ns3d/state.py:13-175)."""

from ...base.state import StatePseudoSpectral


class StateNS3D(StatePseudoSpectral):
    @classmethod
    def _complete_info_solver(cls, info_solver):
        keys_state_phys = ["vx", "vy", "vz"]
        info_solver.classes.State._set_attribs(
            {
                "keys_state_spect": [k + "_fft" for k in keys_state_phys],
                "keys_state_phys": keys_state_phys,
                "keys_phys_needed": keys_state_phys,
                "keys_computable": ["rotz", "divh", "rotz_fft", "divh_fft",
                                    "vt", "vt_fft"],
                "keys_linear_eigenmodes": ["vp_fft", "vt_fft"],
            }
        )

    def compute(self, key, SAVE_IN_DICT=True, RAISE_ERROR=True):
        oper = self.oper
        g = self.state_spect.get_var
        if key == "rotz_fft":
            return 1j * (oper._kdev()[0] * g("vy_fft") - oper._kdev()[1] * g("vx_fft"))
        if key == "rotz":
            return oper.ifft3d(self.compute("rotz_fft").contiguous())
        if key == "divh_fft":
            return 1j * (oper._kdev()[0] * g("vx_fft") + oper._kdev()[1] * g("vy_fft"))
        if key == "divh":
            return oper.ifft3d(self.compute("divh_fft").contiguous())
        if key == "vt_fft":
            return oper.vtfft_from_vecfft(g("vx_fft"), g("vy_fft"), g("vz_fft"))
        if key == "vt":
            return oper.ifft3d(self.compute("vt_fft").contiguous())
        return super().compute(key, SAVE_IN_DICT, RAISE_ERROR)

    def init_from_vxvyvzfft(self, vx_fft, vy_fft, vz_fft):
        """state.py:126-132: set, then a physical round trip."""
        self.state_spect.initialize(0.0)
        self.state_spect.set_var("vx_fft", vx_fft)
        self.state_spect.set_var("vy_fft", vy_fft)
        self.state_spect.set_var("vz_fft", vz_fft)
        self.statephys_from_statespect()
        self.statespect_from_statephys()
        self.invalidate_phys()

    def init_statespect_from(self, **kwargs):
        if len(kwargs) == 1:
            key, field = next(iter(kwargs.items()))
            if key == "vt_fft":
                vx, vy, vz = self.oper.vecfft_from_vtfft(self.oper._to_dev(field, True))
                return super().init_statespect_from(vx_fft=vx, vy_fft=vy, vz_fft=vz)
        return super().init_statespect_from(**kwargs)

    def compute_energy_phys(self):
        ph = self.state_phys
        vx, vy, vz = ph.get_var("vx"), ph.get_var("vy"), ph.get_var("vz")
        return 0.5 * self.oper.mean_space(vx**2 + vy**2 + vz**2)
