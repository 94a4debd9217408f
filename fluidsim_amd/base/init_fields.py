"""Initial fields (fluidsim/base/init_fields.py:26-127, 435-443).

Types: "in_script" (the state stays zero until the script calls
``sim.state.init_statespect_from(...)``), "constant", and the
solver-specific ones registered by subclasses (e.g. "noise").
"from_file" (restart from a state file) is not implemented yet.
"""

from ..params import Parameters


class InitFieldsBase:
    available_types = ("in_script", "constant")

    @classmethod
    def _complete_params_with_default(cls, params, info_solver=None):
        params._set_child(
            "init_fields",
            attribs={
                "type": "constant",
                "available_types": list(cls.available_types),
                "path_file": "",
                "modif_after_init": False,
            },
        )
        params.init_fields._set_child("constant", attribs={"value": 1.0})
        cls._complete_params_specific(params)

    @classmethod
    def _complete_params_specific(cls, params):
        pass

    def __init__(self, sim):
        self.sim = sim
        self.params = sim.params

    def __call__(self):
        t = self.params.init_fields.type
        if t == "in_script":
            self.sim.state.init_statespect_from()
            return
        if t == "constant":
            self._init_constant()
            return
        meth = getattr(self, "init_" + t, None)
        if meth is None:
            raise NotImplementedError(
                f'init_fields.type "{t}" is not available in the MI355X engine '
                f"(available: {self.available_types})"
            )
        meth()

    def _init_constant(self):
        """Every physical variable equals params.init_fields.constant.value:
        only the mean mode is non-zero."""
        value = self.params.init_fields.constant.value
        state = self.sim.state
        state.state_spect.initialize(0.0)
        for ik in range(state.state_spect.nvar):
            state.state_spect.data[ik].view(-1)[0] = value
        state.invalidate_phys()
