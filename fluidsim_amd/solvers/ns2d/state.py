"""State of the ns2d solver (fluidsim/solvers/ns2d/state.py:17-178)."""

from ...base.state import StatePseudoSpectral


class StateNS2D(StatePseudoSpectral):
    @staticmethod
    def _complete_info_solver(info_solver):
        info_solver.classes.State._set_attribs(
            {
                "keys_state_spect": ["rot_fft"],
                "keys_state_phys": ["ux", "uy", "rot"],
                "keys_computable": ["div", "q", "v"],
                "keys_phys_needed": ["rot"],
                "keys_linear_eigenmodes": ["rot_fft"],
            }
        )

    def compute(self, key, SAVE_IN_DICT=True, RAISE_ERROR=True):
        oper = self.oper
        if key == "ux_fft":
            return oper.vecfft_from_rotfft(self.state_spect.get_var("rot_fft"))[0]
        if key == "uy_fft":
            return oper.vecfft_from_rotfft(self.state_spect.get_var("rot_fft"))[1]
        if key == "rot_fft":
            return self.state_spect.get_var("rot_fft")
        if key == "div_fft":
            ux_fft, uy_fft = oper.vecfft_from_rotfft(self.state_spect.get_var("rot_fft"))
            return oper.divfft_from_vecfft(ux_fft, uy_fft)
        if key == "div":
            return oper.ifft2(self.compute("div_fft"))
        if key == "q":
            return self.get_var("rot")
        if key == "v":
            ux = self.state_phys.get_var("ux")
            uy = self.state_phys.get_var("uy")
            return (ux**2 + uy**2).sqrt()
        return super().compute(key, SAVE_IN_DICT, RAISE_ERROR)

    def statephys_from_statespect(self):
        """state.py:95-106: rot, ux, uy from rot_fft."""
        oper = self.oper
        rot_fft = self.state_spect.get_var("rot_fft")
        ux_fft, uy_fft = oper.vecfft_from_rotfft(rot_fft)
        ph = self._state_phys
        oper.ifft_as_arg(rot_fft, ph.get_var("rot"))
        oper.ifft_as_arg(ux_fft.contiguous(), ph.get_var("ux"))
        oper.ifft_as_arg(uy_fft.contiguous(), ph.get_var("uy"))
        self._phys_valid = True

    def statespect_from_statephys(self):
        rot = self._state_phys.get_var("rot")
        self.oper.fft_as_arg(rot, self.state_spect.get_var("rot_fft"))
        self.vars_computed.clear()

    def init_from_rotfft(self, rot_fft):
        """state.py:115-119: dealias, set, refresh the physical state."""
        self.state_spect.initialize(0.0)
        self.state_spect.set_var("rot_fft", rot_fft)
        self.oper.dealiasing(self.state_spect)
        self.invalidate_phys()

    def init_from_uxfft(self, ux_fft):
        uy_fft = self.oper.create_arrayK(value=0.0)
        self.init_from_rotfft(self.oper.rotfft_from_vecfft(self.oper._to_dev(ux_fft, True), uy_fft))

    def init_from_uyfft(self, uy_fft):
        ux_fft = self.oper.create_arrayK(value=0.0)
        self.init_from_rotfft(self.oper.rotfft_from_vecfft(ux_fft, self.oper._to_dev(uy_fft, True)))

    def init_statespect_from(self, **kwargs):
        if len(kwargs) == 1:
            key, arr = next(iter(kwargs.items()))
            if key == "rot_fft":
                return self.init_from_rotfft(arr)
            if key == "ux_fft":
                return self.init_from_uxfft(arr)
            if key == "uy_fft":
                return self.init_from_uyfft(arr)
        return super().init_statespect_from(**kwargs)

    def compute_energy_phys(self):
        ux = self.state_phys.get_var("ux")
        uy = self.state_phys.get_var("uy")
        return 0.5 * self.oper.mean_space(ux**2 + uy**2)
