"""Device-resident SetOfVariables.

Same interface as fluiddyn's SetOfVariables as fluidsim uses it
(fluidsim/base/setofvariables.py:12, fluidsim/base/state.py:243-248): an
array of shape (nvar, *shape_variable) with named variables, ``get_var``,
``set_var``, ``initialize``, ``like=``.  The data is a torch tensor in HBM
(``.data``); ``get_var`` returns a view.  ``to_numpy()`` copies to the host.
"""

import numpy as np


def _torch():
    import torch

    return torch


class SetOfVariables:
    def __init__(self, keys=None, shape_variable=None, dtype=None, info=None,
                 like=None, value=None, device="cuda"):
        torch = _torch()
        if like is not None:
            self.keys = list(like.keys)
            self.info = like.info if info is None else info
            self.data = torch.zeros_like(like.data)
        else:
            if keys is None or shape_variable is None or dtype is None:
                raise ValueError("keys, shape_variable and dtype are required")
            self.keys = list(keys)
            self.info = info
            self.data = torch.zeros(
                (len(self.keys),) + tuple(shape_variable), dtype=dtype,
                device=device,
            )
        if value is not None:
            self.data.fill_(value)

    @property
    def nvar(self):
        return len(self.keys)

    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def dtype(self):
        return self.data.dtype

    def __len__(self):
        return self.nvar

    def __getitem__(self, i):
        return self.data[i]

    def __setitem__(self, i, value):
        self.data[i] = _as_tensor(value, self.data)

    def get_var(self, key):
        return self.data[self.keys.index(key)]

    def set_var(self, key, value):
        self.data[self.keys.index(key)] = _as_tensor(value, self.data)

    def initialize(self, value=0.0):
        self.data.fill_(value)

    def fill(self, value):
        self.data.fill_(value)

    def to_numpy(self):
        return self.data.detach().cpu().numpy()

    def __iadd__(self, other):
        self.data += other.data if isinstance(other, SetOfVariables) else other
        return self

    def __repr__(self):
        return f"SetOfVariables(keys={self.keys}, shape={self.shape}, dtype={self.dtype})"


def _as_tensor(value, like):
    torch = _torch()
    if isinstance(value, torch.Tensor):
        return value.to(device=like.device, dtype=like.dtype)
    arr = np.asarray(value)
    t = torch.from_numpy(np.ascontiguousarray(arr))
    return t.to(device=like.device, dtype=like.dtype)
