"""Extension of a Simul class with "Simul extenders"
(lib/fluidsim_core/extend_simul.py:24-75).

``extend_simul_class(Simul, extenders)`` returns a subclass whose
InfoSolver applies each extender's ``get_modif_info_solver()`` function
after the default classes are registered; extenders also complete the
parameters (``complete_params_with_default``)."""

from abc import ABCMeta, abstractmethod


def extend_simul_class(Simul, extenders):
    if isinstance(extenders, type):
        extenders = [extenders]

    class NewInfoSolver(Simul.InfoSolver):
        pass

    NewInfoSolver._modificators = list(getattr(Simul.InfoSolver, "_modificators", []))
    NewInfoSolver._extenders = list(getattr(Simul.InfoSolver, "_extenders", []))
    for extender in extenders:
        NewInfoSolver._modificators.append(extender.get_modif_info_solver())
        NewInfoSolver._extenders.append(extender)

    class NewSimul(Simul):
        InfoSolver = NewInfoSolver

    NewSimul.__name__ = Simul.__name__
    return NewSimul


class SimulExtender(metaclass=ABCMeta):
    """Base of the classes that extend a Simul class
    (lib/fluidsim_core/extend_simul.py:40-75)."""

    @classmethod
    @abstractmethod
    def get_modif_info_solver(cls):
        """Return a function that modifies ``info_solver``."""

    @classmethod
    @abstractmethod
    def complete_params_with_default(cls, params):
        """Complete the simulation parameters."""

    @classmethod
    def _complete_params_with_default(cls, params):
        cls.complete_params_with_default(params)
