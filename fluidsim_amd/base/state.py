"""Device-resident state (fluidsim/base/state.py:209-395).

``state_spect`` lives in HBM and is advanced in place by the engine.
``state_phys`` is refreshed lazily: the reference recomputes it after every
step (one_time_step_computation, fluidsim/base/time_stepping/pseudo_spect.py
:236-243); here it is recomputed when read after the spectral state changed,
which gives the same values without the per-step inverse transforms.
"""

import numpy as np

from ..setofvariables import SetOfVariables


class StatePseudoSpectral:
    @staticmethod
    def _complete_info_solver(info_solver):
        info_solver.classes.State._set_attribs(
            {
                "keys_state_spect": ["ux_fft", "uy_fft"],
                "keys_state_phys": ["ux", "uy"],
                "keys_computable": [],
                "keys_phys_needed": ["ux", "uy"],
                "keys_linear_eigenmodes": [],
            }
        )

    def __init__(self, sim, oper=None):
        self.sim = sim
        self.params = sim.params
        self.oper = sim.oper if oper is None else oper
        info = sim.info_solver.classes.State
        self.keys_state_spect = list(info.keys_state_spect)
        self.keys_state_phys = list(info.keys_state_phys)
        self.keys_computable = list(info.keys_computable)
        self.keys_phys_needed = list(info.keys_phys_needed)
        self.state_spect = SetOfVariables(
            keys=self.keys_state_spect, shape_variable=self.oper.shapeK,
            dtype=self.oper.dtype_complex, info="state_spect",
        )
        self._state_phys = SetOfVariables(
            keys=self.keys_state_phys, shape_variable=self.oper.shapeX,
            dtype=self.oper.dtype_real, info="state_phys",
        )
        self._phys_valid = False
        self.vars_computed = {}
        self.it_computed = {}

    # --- physical state -------------------------------------------------
    @property
    def state_phys(self):
        if not self._phys_valid:
            self.statephys_from_statespect()
        return self._state_phys

    def invalidate_phys(self):
        self._phys_valid = False
        self.vars_computed.clear()
        self.it_computed.clear()

    def statephys_from_statespect(self):
        for ik in range(self.state_spect.nvar):
            self.oper.ifft_as_arg(self.state_spect[ik], self._state_phys[ik])
        self._phys_valid = True

    def statespect_from_statephys(self):
        for ik in range(self.state_spect.nvar):
            self.oper.fft_as_arg(self._state_phys[ik], self.state_spect[ik])
        self.vars_computed.clear()

    def has_vars(self, *keys):
        known = set(self.keys_state_phys + self.keys_computable + self.keys_state_spect)
        return set(keys).issubset(known)

    def can_this_key_be_obtained(self, key):
        return self.has_vars(key)

    def _it(self):
        ts = getattr(self.sim, "time_stepping", None)
        return ts.it if ts is not None else 0

    def get_var(self, key):
        """Device tensor of a variable (a view for stored variables)."""
        if key in self.keys_state_spect:
            return self.state_spect.get_var(key)
        if key in self.keys_state_phys:
            return self.state_phys.get_var(key)
        it = self._it()
        if key in self.vars_computed and it == self.it_computed[key]:
            return self.vars_computed[key]
        value = self.compute(key)
        self.vars_computed[key] = value
        self.it_computed[key] = it
        return value

    def get_var_numpy(self, key):
        """Host (numpy) copy of a variable."""
        return self.get_var(key).detach().cpu().numpy()

    def compute(self, key, SAVE_IN_DICT=True, RAISE_ERROR=True):
        if RAISE_ERROR:
            raise ValueError(f'Do not know how to compute "{key}".')
        return self.oper.create_arrayX(value=0.0)

    def __setitem__(self, key, value):
        if key in self.keys_state_spect:
            self.state_spect.set_var(key, value)
            self.invalidate_phys()
        elif key in self.keys_state_phys:
            self.state_phys.set_var(key, value)
        else:
            raise ValueError('key "' + key + '" is not known')

    def init_statespect_from(self, **kwargs):
        """fluidsim/base/state.py:352-383: other keys are set to zero."""
        self.state_spect.initialize(0.0)
        for key, value in kwargs.items():
            if key not in self.keys_state_spect:
                raise ValueError(
                    f"Do not know how to initialize with key '{key}'. "
                    f"({self.keys_state_spect = })"
                )
            self.state_spect.set_var(key, value)
        self.invalidate_phys()

    def init_statephys_from(self, **kwargs):
        self._state_phys.initialize(0.0)
        for key, value in kwargs.items():
            if key not in self.keys_state_phys:
                raise ValueError(f'Do not know how to initialize with key "{key}".')
            self._state_phys.set_var(key, value)
        self._phys_valid = True

    def check_energy_equal_phys_spect(self):
        """fluidsim/base/state.py:385-392."""
        energy_spect = self.sim.output.compute_energy()
        energy_phys = self.compute_energy_phys()
        if not np.allclose(energy_spect, energy_phys):
            raise RuntimeError(
                "Physical and spectral states are inconsistent: "
                f"{energy_spect} != {energy_phys}"
            )

    def compute_energy_phys(self):
        raise NotImplementedError
