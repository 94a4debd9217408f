"""Forcing organiser (fluidsim/base/forcing/base.py:26-221).

``ForcingBasePseudoSpectral`` completes ``params.forcing`` from the forcing
makers registered in ``info_solver.classes.Forcing.classes``, builds the one
selected by ``params.forcing.type`` and forwards ``compute`` (once per time
step, after the CFL update) and ``get_forcing``.
"""

import importlib

import numpy as np

from .specific import (
    InScriptForcingPseudoSpectralCoarse,
    TimeCorrelatedRandomPseudoSpectral,
    TimeCorrelatedRandomPseudoSpectralAnisotropic,
    TimeCorrelatedRandomPseudoSpectralAnisotropic3D,
)


def _maker_classes(info_solver):
    node = info_solver.classes.Forcing.classes
    out = {}
    for tag in node._tag_children:
        child = getattr(node, tag)
        mod = importlib.import_module(child.module_name)
        out[tag] = getattr(mod, child.class_name)
    return out


class ForcingBase:
    _name_task = "forcing"

    @staticmethod
    def _complete_info_solver(info_solver, classes=None):
        """base.py:31-45: one child per forcing maker under
        classes.Forcing.classes."""
        node = info_solver.classes.Forcing._set_child("classes")
        for cls in classes or ():
            node._set_child(
                cls.tag,
                attribs={"module_name": cls.__module__, "class_name": cls.__name__},
            )

    @staticmethod
    def _complete_params_with_default(params, info_solver):
        """base.py:47-78."""
        params._set_child(
            "forcing",
            attribs={
                "enable": False,
                "type": "",
                "available_types": [],
                "forcing_rate": 1.0,
                "key_forced": None,
            },
        )
        for tag, cls in sorted(_maker_classes(info_solver).items()):
            cls._complete_params_with_default(params)

    def __init__(self, sim):
        """base.py:100-132."""
        params = sim.params
        self.type_forcing = params.forcing.type
        dict_classes = _maker_classes(sim.info_solver)
        if self.type_forcing not in dict_classes:
            known = ", ".join(sorted(dict_classes)) or "none registered"
            raise ValueError(
                f"params.forcing.type = {self.type_forcing!r} names no forcing of this "
                f"solver (known: {known})"
            )
        self.sim = sim
        self.forcing_maker = dict_classes[self.type_forcing](sim)
        self._t_last_computed = -np.inf

    def __call__(self, key):
        return self.get_forcing().get_var(key)

    def compute(self):
        """base.py:140-144: once per simulation time."""
        time = self.sim.time_stepping.t
        if time > self._t_last_computed:
            self.forcing_maker.compute()
            self._t_last_computed = time

    def get_forcing(self):
        return self.forcing_maker.forcing_fft

    def is_initialized(self):
        return getattr(self.forcing_maker, "is_initialized", True)


class ForcingBasePseudoSpectral(ForcingBase):
    """base.py:170-221."""

    @staticmethod
    def _complete_info_solver(info_solver, classes=None):
        defaults = [
            InScriptForcingPseudoSpectralCoarse,
            TimeCorrelatedRandomPseudoSpectral,
        ]
        classes = list(classes or [])
        for c in defaults:
            if not any(c.tag == k.tag for k in classes):
                classes.append(c)
        ForcingBase._complete_info_solver(info_solver, classes)

    @staticmethod
    def _complete_params_with_default(params, info_solver):
        ForcingBase._complete_params_with_default(params, info_solver)
        params.forcing._set_attribs({"nkmax_forcing": 5, "nkmin_forcing": 4})


class ForcingNS2D(ForcingBasePseudoSpectral):
    """fluidsim/solvers/ns2d/forcing.py:23-64."""

    @staticmethod
    def _complete_info_solver(info_solver, classes=None):
        ForcingBasePseudoSpectral._complete_info_solver(
            info_solver, [TimeCorrelatedRandomPseudoSpectralAnisotropic]
        )


class ForcingNS3D(ForcingBasePseudoSpectral):
    """fluidsim/solvers/ns3d/forcing/__init__.py:86-101 (the tcrandom
    makers; taylor_green, milestone and watu are not provided)."""

    @staticmethod
    def _complete_info_solver(info_solver, classes=None):
        ForcingBasePseudoSpectral._complete_info_solver(
            info_solver, [TimeCorrelatedRandomPseudoSpectralAnisotropic3D]
        )
