"""Base solver classes.

SimulBasePseudoSpectral mirrors fluidsim's SimulBase /
SimulBasePseudoSpectral (fluidsim/base/solvers/base.py:117-223,
fluidsim/base/solvers/pseudo_spect.py:101-191): the same parameter
completion chain and construction order (Operators, Output, State,
TimeStepping, InitFields, then output.post_init()).
"""

import inspect

import numpy as np

from ..info import InfoSolver
from ..params import Parameters


class InfoSolverPseudoSpectral(InfoSolver):
    """Classes of a 2D pseudo-spectral solver
    (fluidsim/base/solvers/pseudo_spect.py:31-73)."""

    def _init_root(self):
        self.module_name = "fluidsim_amd.base.solvers"
        self.class_name = "SimulBasePseudoSpectral"
        self.short_name = "BasePS"
        self._set_class("Operators", "fluidsim_amd.operators.operators2d",
                        "OperatorsPseudoSpectral2D")
        self._set_class("State", "fluidsim_amd.base.state", "StatePseudoSpectral")
        self._set_class("TimeStepping", "fluidsim_amd.base.time_stepping",
                        "TimeSteppingPseudoSpectral")
        self._set_class("InitFields", "fluidsim_amd.base.init_fields",
                        "InitFieldsBase")
        self._set_class("Output", "fluidsim_amd.base.output",
                        "OutputBasePseudoSpectral")
        self._set_class("Forcing", "fluidsim_amd.base.forcing",
                        "ForcingBasePseudoSpectral")


class InfoSolverPseudoSpectral3D(InfoSolverPseudoSpectral):
    def _init_root(self):
        super()._init_root()
        self.classes.Operators.module_name = "fluidsim_amd.operators.operators3d"
        self.classes.Operators.class_name = "OperatorsPseudoSpectral3D"


def _complete(cls, params, info_solver):
    f = getattr(cls, "_complete_params_with_default", None)
    if f is None:
        return
    nargs = len(inspect.signature(f).parameters)
    if nargs >= 2:
        f(params, info_solver)
    else:
        f(params)


class SimulBasePseudoSpectral:
    InfoSolver = InfoSolverPseudoSpectral

    @classmethod
    def create_default_params(cls):
        """lib/fluidsim_core/solver.py:33-40 and params.py:22-78."""
        info_solver = cls.InfoSolver()
        info_solver.complete_with_classes()
        params = Parameters(tag="params")
        cls._complete_params_with_default(params)
        for role, klass in info_solver.import_classes().items():
            _complete(klass, params, info_solver)
        return params

    @staticmethod
    def _complete_params_with_default(params):
        params._set_attribs(
            {
                "short_name_type_run": "",
                "NEW_DIR_RESULTS": True,
                "ONLY_COARSE_OPER": False,
                "nu_2": 0.0,
                "nu_4": 0.0,
                "nu_8": 0.0,
                "nu_m4": 0.0,
            }
        )

    def __init__(self, params):
        self.params = params
        self.info_solver = self.InfoSolver()
        self.info_solver.complete_with_classes()
        dict_classes = self.info_solver.import_classes()
        self.oper = dict_classes["Operators"](params=params)
        self.output = dict_classes["Output"](self)
        self.state = dict_classes["State"](self)
        self.time_stepping = dict_classes["TimeStepping"](self)
        self.init_fields = dict_classes["InitFields"](self)
        self.init_fields()
        # other classes with a params.<task>.enable switch
        # (fluidsim/base/solvers/base.py:190-213): the forcing
        simple = ("Operators", "Output", "State", "TimeStepping", "InitFields")
        for role, cls in dict_classes.items():
            if role in simple:
                continue
            name_task = cls._name_task
            setattr(self, f"is_{name_task}_enabled", False)
            params_cls = getattr(params, name_task, None)
            if params_cls is not None and params_cls.enable:
                setattr(self, f"is_{name_task}_enabled", True)
                setattr(self, name_task, cls(self))
        self.output.post_init()

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.output.end_of_simul()

    @property
    def name_run(self):
        return self.output.name_run

    def compute_freq_diss(self):
        """fluidsim/base/solvers/pseudo_spect.py:134-191 (host arrays)."""
        p = self.params
        oper = self.oper
        f_d = p.nu_2 * oper.K2 if p.nu_2 > 0 else np.zeros(oper.shapeK)
        if p.nu_4 > 0.0:
            f_d = f_d + p.nu_4 * oper.K4
        if p.nu_8 > 0.0:
            f_d = f_d + p.nu_8 * oper.K8
        if p.nu_m4 != 0.0:
            f_d_hypo = p.nu_m4 / oper.K2_not0**2
            f_d_hypo.flat[0] = f_d_hypo.flat[1]
        else:
            f_d_hypo = 0.0
        return f_d, f_d_hypo

    def visc_coefs(self):
        p = self.params
        return [float(p.nu_2), float(p.nu_4), float(p.nu_8), float(p.nu_m4)]

    def tendencies_nonlin(self, state_spect=None, old=None):
        from ..setofvariables import SetOfVariables

        tend = SetOfVariables(like=self.state.state_spect) if old is None else old
        tend.initialize(0.0)
        return tend


Simul = SimulBasePseudoSpectral
