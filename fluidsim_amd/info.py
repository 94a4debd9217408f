"""Solver information and class registry.

Mirrors fluidsim_core's InfoSolverCore (lib/fluidsim_core/info.py:103-126):
``classes.<Role>.module_name / class_name`` strings are imported by
``import_classes``, so a user can swap any role (Operators, State,
TimeStepping, InitFields, Output) by editing the strings before
``Simul(params)``, as in fluidsim.
"""

import importlib

from .params import Parameters


class InfoSolver(Parameters):
    def __init__(self):
        super().__init__(tag="solver")
        self._set_attribs(
            {"module_name": None, "class_name": None, "short_name": None}
        )
        self._set_child("classes")
        self._init_root()

    def _init_root(self):
        pass

    def _set_class(self, role, module_name, class_name):
        self.classes._set_child(
            role, attribs={"module_name": module_name, "class_name": class_name}
        )

    def import_classes(self):
        out = {}
        for role in self.classes._tag_children:
            child = getattr(self.classes, role)
            mod = importlib.import_module(child.module_name)
            out[role] = getattr(mod, child.class_name)
        return out

    def complete_with_classes(self):
        """Let each class add its info (fluidsim's _complete_info_solver),
        then apply the modifications of the Simul extenders
        (lib/fluidsim_core/info.py:60-75)."""
        for cls in self.import_classes().values():
            f = getattr(cls, "_complete_info_solver", None)
            if f is not None:
                f(self)
        for modif in getattr(self, "_modificators", ()):
            modif(self)
