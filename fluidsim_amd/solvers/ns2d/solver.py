"""NS2D solver on the MI355X engine.

Drop-in for fluidsim.solvers.ns2d.solver (fluidsim/solvers/ns2d/solver.py):
2D incompressible Navier-Stokes in vorticity form,
d_t rot_hat = N(rot) - sigma(k) rot_hat with N = -u.grad(rot) - beta v.
"""

from ... import _lib
from ...base.solvers import SimulBasePseudoSpectral, InfoSolverPseudoSpectral
from ...setofvariables import SetOfVariables


class InfoSolverNS2D(InfoSolverPseudoSpectral):
    """fluidsim/solvers/ns2d/solver.py:42-92."""

    def _init_root(self):
        super()._init_root()
        package = "fluidsim_amd.solvers.ns2d"
        self.module_name = package + ".solver"
        self.class_name = "Simul"
        self.short_name = "NS2D"
        self._set_class("State", package + ".state", "StateNS2D")
        self._set_class("InitFields", package + ".init_fields", "InitFieldsNS2D")
        self._set_class("Output", package + ".output", "OutputNS2D")
        self._set_class("Forcing", "fluidsim_amd.base.forcing", "ForcingNS2D")


class Simul(SimulBasePseudoSpectral):
    InfoSolver = InfoSolverNS2D
    _engine_projection = 0
    _cfl_ncomp = 2

    @staticmethod
    def _complete_params_with_default(params):
        SimulBasePseudoSpectral._complete_params_with_default(params)
        params._set_attribs({"beta": 0.0})

    def _engine_tendency(self, rot_fft, out_fft, want_maxima=False):
        """Fused device tendency of one vorticity field (tensors (1, *K))."""
        self.oper.ns2d_tendency(rot_fft, out_fft, self.params.beta, want_maxima)

    def _engine_rk_stage(self, scheme, stage, S, tmp, t1, visc, cfl, dt_max):
        """One fused Runge-Kutta stage on the device (fsx_ns2d_rk_stage)."""
        _lib.call("fsx_ns2d_rk_stage", self.oper.plan.handle, scheme, stage,
                  _lib.dptr(S), _lib.dptr(tmp), _lib.dptr(t1),
                  float(self.params.beta), visc, cfl, dt_max, _lib.stream_ptr())

    def tendencies_nonlin(self, state_spect=None, old=None):
        """fluidsim/solvers/ns2d/solver.py:111-194.

        Returns a SetOfVariables with the dealiased N(rot); ``old`` (if
        given) is reused as output storage."""
        if state_spect is None:
            state_spect = self.state.state_spect
        tend = SetOfVariables(like=self.state.state_spect) if old is None else old
        if old is not None:
            # modes outside the kept box are never written by the engine
            tend.initialize(0.0)
        self._engine_tendency(state_spect.data, tend.data)
        if self.is_forcing_enabled:
            # solver.py:190-191, after the dealiasing
            _lib.call("fsx_forcing_add", self.oper.plan.handle, _lib.dptr(tend.data),
                      1, 1.0, _lib.stream_ptr())
        return tend
