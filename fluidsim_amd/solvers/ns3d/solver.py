"""NS3D solver on the MI355X engine.

Drop-in for fluidsim.solvers.ns3d.solver (fluidsim/solvers/ns3d/solver.py):
incompressible Navier-Stokes in rotational form,
N(v) = P(fft(v x omega)) with omega = curl v (+ f e_z at k = 0), P the Leray
projection (or toroidal / poloidal with params.projection).
"""

from ... import _lib
from ...base.solvers import SimulBasePseudoSpectral, InfoSolverPseudoSpectral3D
from ...operators.operators3d import PROJ
from ...setofvariables import SetOfVariables


class InfoSolverNS3D(InfoSolverPseudoSpectral3D):
    """fluidsim/solvers/ns3d/solver.py:31-54."""

    def _init_root(self):
        super()._init_root()
        package = "fluidsim_amd.solvers.ns3d"
        self.module_name = package + ".solver"
        self.class_name = "Simul"
        self.short_name = "ns3d"
        self._set_class("State", package + ".state", "StateNS3D")
        self._set_class("InitFields", package + ".init_fields", "InitFieldsNS3D")
        self._set_class("Output", package + ".output", "OutputNS3D")
        self._set_class("Forcing", "fluidsim_amd.base.forcing", "ForcingNS3D")


class Simul(SimulBasePseudoSpectral):
    InfoSolver = InfoSolverNS3D
    _cfl_ncomp = 3

    @staticmethod
    def _complete_params_with_default(params):
        SimulBasePseudoSpectral._complete_params_with_default(params)
        params._set_attribs({"f": None, "no_vz_kz0": False, "projection": None})

    def __init__(self, params):
        self._init_projection_params(params)
        super().__init__(params)

    def _init_projection_params(self, params):
        """_init_projection (solver.py:147-174)."""
        self.no_vz_kz0 = bool(getattr(params, "no_vz_kz0", False))
        projection = getattr(params, "projection", None)
        if projection not in (None, "toroidal", "vortical", "poloidal"):
            raise ValueError(
                "No known projection for params.projection = "
                f"{params.projection}"
            )
        self._engine_projection = PROJ[projection] if projection else 1

    def _init_projection(self):
        self._init_projection_params(self.params)

    def _engine_tendency(self, v_fft, out_fft, want_maxima=False):
        """Fused fft(v x omega), dealiased, unprojected (projection is
        folded into the Runge-Kutta stage)."""
        self.oper.ns3d_tendency(v_fft, out_fft, f=self.params.f,
                                want_maxima=want_maxima)
        if getattr(self, "is_turb_model_enabled", False):
            # solver.py:246-249: model term of the same state
            self.turb_model._model.add_to(v_fft, out_fft)

    def _engine_rk_stage(self, scheme, stage, S, tmp, t1, visc, cfl, dt_max):
        """One fused Runge-Kutta stage on the device (fsx_ns3d_rk_stage)."""
        f = self.params.f
        _lib.call("fsx_ns3d_rk_stage", self.oper.plan.handle, scheme, stage,
                  _lib.dptr(S), _lib.dptr(tmp), _lib.dptr(t1), int(f is not None),
                  float(f) if f is not None else 0.0, self._engine_projection,
                  int(self.no_vz_kz0), visc, cfl, dt_max, _lib.stream_ptr())

    def project_state_spect(self, state_spect):
        """solver.py:255-263."""
        d = state_spect.data if isinstance(state_spect, SetOfVariables) else state_spect
        kind = self._engine_projection
        proj = {1: self.oper.project_perpk3d, 2: self.oper.project_toroidal,
                3: self.oper.project_poloidal}[kind]
        proj(d[0], d[1], d[2])
        if self.no_vz_kz0:
            d[2][0] = 0.0
            if d.shape[0] > 3:
                d[3][0] = 0.0

    def tendencies_nonlin(self, state_spect=None, old=None):
        """solver.py:180-253 (no forcing / turb_model)."""
        if state_spect is None:
            state_spect = self.state.state_spect
        tend = SetOfVariables(like=self.state.state_spect) if old is None else old
        if old is not None:
            tend.initialize(0.0)
        self._engine_tendency(state_spect.data, tend.data)
        if self.is_forcing_enabled:
            # solver.py:243-244, before the projection
            _lib.call("fsx_forcing_add", self.oper.plan.handle, _lib.dptr(tend.data),
                      tend.nvar, 1.0, _lib.stream_ptr())
        self.project_state_spect(tend)
        return tend
