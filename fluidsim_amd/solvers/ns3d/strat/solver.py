"""NS3D.strat solver on the MI355X engine.

Drop-in for fluidsim.solvers.ns3d.strat.solver
(fluidsim/solvers/ns3d/strat/solver.py): Boussinesq equations with a constant
Brunt-Vaisala frequency N,
  d_t v = P(v x omega + b e_z) - sigma v,   d_t b = -v.grad b - N^2 v_z - sigma b
(one sigma(k) for all variables, fluidsim/base/time_stepping/pseudo_spect.py
:179-189).
"""

from .... import _lib
from ....setofvariables import SetOfVariables
from ..solver import InfoSolverNS3D, Simul as SimulNS3D


class InfoSolverNS3DStrat(InfoSolverNS3D):
    """strat/solver.py:36-59."""

    def _init_root(self):
        super()._init_root()
        package = "fluidsim_amd.solvers.ns3d.strat"
        self.module_name = package + ".solver"
        self.class_name = "Simul"
        self.short_name = "ns3d.strat"
        self._set_class("State", package + ".state", "StateNS3DStrat")
        self._set_class("Output", package + ".output", "OutputNS3DStrat")


class Simul(SimulNS3D):
    InfoSolver = InfoSolverNS3DStrat
    # strat/solver.py:138-216 adds no turbulence-model term
    _turb_model_in_tendency = False

    @staticmethod
    def _complete_params_with_default(params):
        SimulNS3D._complete_params_with_default(params)
        params._set_attribs({"N": 1.0})

    def _engine_tendency(self, state, out_fft, want_maxima=False):
        """Strat tendency (4 fields), projected and dealiased."""
        f = self.params.f
        _lib.call("fsx_strat_tendency", self.oper.plan.handle, _lib.dptr(state),
                  _lib.dptr(out_fft), int(f is not None),
                  float(f) if f is not None else 0.0, float(self.params.N),
                  self._engine_projection, int(self.no_vz_kz0), int(want_maxima),
                  _lib.stream_ptr())

    def _engine_rk_stage(self, scheme, stage, S, tmp, t1, visc, cfl, dt_max):
        f = self.params.f
        _lib.call("fsx_strat_rk_stage", self.oper.plan.handle, scheme, stage,
                  _lib.dptr(S), _lib.dptr(tmp), _lib.dptr(t1), int(f is not None),
                  float(f) if f is not None else 0.0, float(self.params.N),
                  self._engine_projection, int(self.no_vz_kz0), visc, cfl, dt_max,
                  _lib.stream_ptr())

    def tendencies_nonlin(self, state_spect=None, old=None):
        """strat/solver.py:138-216."""
        if state_spect is None:
            state_spect = self.state.state_spect
        tend = SetOfVariables(like=self.state.state_spect) if old is None else old
        if old is not None:
            tend.initialize(0.0)
        self._engine_tendency(state_spect.data, tend.data)
        if self.is_forcing_enabled:
            # strat/solver.py:209-213: forcing added before the projection
            # (the engine tendency is projected: P(P N + F) = P(N + F))
            _lib.call("fsx_forcing_add", self.oper.plan.handle, _lib.dptr(tend.data),
                      tend.nvar, 1.0, _lib.stream_ptr())
            self.project_state_spect(tend)
        return tend
