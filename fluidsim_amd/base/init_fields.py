"""Initial fields (fluidsim/base/init_fields.py:26-127, 140-297, 435-443).

Types: "in_script" (the state stays zero until the script calls
``sim.state.init_statespect_from(...)``), "constant", "from_simul"
(``sim.init_fields.get_state_from_simul(sim_in)`` copies the state of
another simulation, changing the resolution on the device;
fluidsim/base/init_fields.py:340-432), "from_file" (restart from a state
file of fluidsim_amd.base.phys_fields, ``params.init_fields.from_file.path``)
and the solver-specific ones registered by subclasses (e.g. "noise").
"""

import numpy as np

from .. import _lib
from ..params import Parameters


class InitFieldsBase:
    available_types = ("in_script", "constant", "from_simul", "from_file")

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
        params.init_fields._set_child("from_file", attribs={"path": ""})
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

    def init_from_file(self):
        """InitFieldsFromFile.__call__ (fluidsim/base/init_fields.py:157-297):
        the file's grid and box must be the simulation's; the physical
        fields the state needs are read (missing ones are zero), the
        spectral state is their transform, and time and iteration are the
        file's."""
        from .phys_fields import StateFile

        sim = self.sim
        path = str(self.params.init_fields.from_file.path)
        f = StateFile(path)
        print("Load state from file:\n[...]" + path[-75:])
        try:
            oper_file = f.params.get("oper", {})
            po = self.params.oper
            for r in f.attrs.get("axes", []):
                nr, Lr = f"n{r}", f"L{r}"
                if po[nr] != oper_file.get(nr):
                    raise ValueError(
                        "this is not a correct state for this simulation\n"
                        "self.{0} != params_file.{0}".format(nr))
                if Lr in oper_file and po[Lr] != oper_file[Lr]:
                    raise ValueError(
                        "this is not a correct state for this simulation\n"
                        "self.params.oper.{0} != params_file.{0}".format(Lr))
            state = sim.state
            state._state_phys.initialize(0.0)
            needed = sim.info_solver.classes.State.keys_phys_needed
            for k in needed:
                if k in f.keys:
                    state._state_phys.set_var(k, np.ascontiguousarray(f[k]))
            t = float(f.state_attrs.get("time", 0.0))
            it = int(f.state_attrs.get("it", 0))
        finally:
            f.close()
        state.statespect_from_statephys()
        state.invalidate_phys()
        sim.time_stepping.t = t
        sim.time_stepping.it = it

    def init_from_simul(self):
        """InitFieldsFromSimul.__call__ (init_fields.py:343-344)."""
        self.get_state_from_simul = self._get_state_from_simul

    def _get_state_from_simul(self, sim_in):
        """init_fields.py:384-432: time and state of sim_in, resolution
        changed with fill_field_fft_2d / fill_field_fft_3d on the device."""
        sim = self.sim
        if sim.output.name_solver != sim_in.output.name_solver:
            raise NotImplementedError
        sim.time_stepping.t = sim_in.time_stepping.t
        src = sim_in.state.state_spect.data
        dst = sim.state.state_spect
        if tuple(src.shape) == tuple(dst.data.shape):
            dst.data.copy_(src.to(dst.data.dtype))
        else:
            if src.dtype != dst.data.dtype:
                src = src.to(dst.data.dtype)
            src = src.contiguous()
            dst.initialize(0.0)
            si, so = tuple(src.shape[1:]), tuple(dst.data.shape[1:])
            if len(si) == 2:
                si, so = (1,) + si, (1,) + so
            _lib.call("fsx_resize_fft", _lib.dptr(src), si[0], si[1], si[2],
                      _lib.dptr(dst.data), so[0], so[1], so[2], src.shape[0],
                      sim.oper.precision, _lib.stream_ptr())
        sim.state.invalidate_phys()

    def _init_constant(self):
        """Every physical variable equals params.init_fields.constant.value:
        only the mean mode is non-zero."""
        value = self.params.init_fields.constant.value
        state = self.sim.state
        state.state_spect.initialize(0.0)
        for ik in range(state.state_spect.nvar):
            state.state_spect.data[ik].view(-1)[0] = value
        state.invalidate_phys()
