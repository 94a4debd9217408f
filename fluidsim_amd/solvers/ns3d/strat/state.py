"""State of the ns3d.strat solver (fluidsim/solvers/ns3d/strat/state.py:8-43)."""

from ..state import StateNS3D


class StateNS3DStrat(StateNS3D):
    @classmethod
    def _complete_info_solver(cls, info_solver):
        StateNS3D._complete_info_solver(info_solver)
        keys_state_phys = ["vx", "vy", "vz", "b"]
        info_solver.classes.State._set_attribs(
            {
                "keys_state_spect": [k + "_fft" for k in keys_state_phys],
                "keys_state_phys": keys_state_phys,
                "keys_phys_needed": keys_state_phys,
            }
        )

    def init_from_vxvyvzfft(self, vx_fft, vy_fft, vz_fft):
        self.state_spect.initialize(0.0)
        super().init_from_vxvyvzfft(vx_fft, vy_fft, vz_fft)
