"""Device diagnostics: spectra, spectral energy budgets and spatial means.

Host side of ``fsx_diag`` (one pass over the K arrays, FP64 histograms in
LDS) and the scheduling of the specific outputs
(fluidsim/base/output/base.py:524-693, SpecificOutput).  Results are kept
in memory and, when ``params.output.HAS_TO_SAVE``, written next to the run:
``<name>.npz`` with a ``times`` array and one stacked array per key (the
reference writes the same datasets to HDF5, fluidsim/base/output/base.py
:612-668; h5py is not available here), and ``spatial_means.txt`` in the
reference's text format.
"""

import ctypes
import os

import numpy as np

from .. import _lib

KX, KY, KZ, SHELL, KZKH = 1, 2, 4, 8, 16

NS2D_SPECTRA, NS2D_BUDGET, NS3D_SPECTRA, STRAT_SPECTRA = 0, 1, 2, 3
NS3D_BUDGET, STRAT_BUDGET, NS2D_MEANS, NS3D_MEANS, STRAT_MEANS = 4, 5, 6, 7, 8


def run(oper, kind, flags, fields, params=(0.0, 0.0, 0.0, 0.0, 0.0)):
    """fsx_diag on device fields (contiguous K tensors of the plan
    precision); returns the histograms as (Q, ...) numpy arrays under the
    keys kx, ky, kz, shell, kzkh and the Parseval sums under 'sums'."""
    plan = oper.plan.handle
    lay = (ctypes.c_longlong * 8)()
    _lib.call("fsx_diag_layout", plan, kind, flags, lay)
    Q, nkx, nky, nkz, nk, nkh, total, _ = (int(v) for v in lay)
    for f in fields:
        if f is not None and not f.is_contiguous():
            raise ValueError("diagnostics need contiguous fields")
    ptrs = (ctypes.c_void_p * len(fields))(
        *[_lib.dptr(f) if f is not None else None for f in fields])
    out = np.empty(total, dtype=np.float64)
    _lib.call("fsx_diag", plan, kind, flags, ptrs, len(fields), _lib.doubles(params),
              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _lib.stream_ptr())
    res = {}
    off = 0
    for flag, key, shape in ((KX, "kx", (Q, nkx)), (KY, "ky", (Q, nky)), (KZ, "kz", (Q, nkz)),
                             (SHELL, "shell", (Q, nk)), (KZKH, "kzkh", (Q, nkz, nkh))):
        if flags & flag:
            size = int(np.prod(shape))
            res[key] = out[off:off + size].reshape(shape)
            off += size
    res["sums"] = out[off:off + Q]
    return res


def diss_params(sim, N=None):
    inv_N2 = 0.0 if N is None else 1.0 / N**2
    return [inv_N2] + list(sim.visc_coefs())


class SpecificOutput:
    """Periodic computation and storage of one output
    (fluidsim/base/output/base.py:524-693)."""

    _tag = ""
    _name_file = None

    def __init__(self, output, period_save=0.0, arrays_1st_time=None):
        self.output = output
        self.sim = output.sim
        self.oper = output.sim.oper
        self.params = output.sim.params
        self.period_save = float(period_save or 0.0)
        self.arrays_1st_time = arrays_1st_time or {}
        self.times = []
        self.records = []
        self.nb_saved_times = 0
        self.t_last_save = None
        self.path_file = (os.path.join(output.path_run, self._name_file)
                          if self._name_file else None)

    # scheduling ------------------------------------------------------
    def _init_files(self):
        if not self.period_save:
            return
        self._save(self.compute())
        self.t_last_save = self.sim.time_stepping.t

    def due(self, tsim):
        return bool(self.period_save) and (
            self.t_last_save is None or tsim - self.t_last_save >= self.period_save)

    def _online_save(self):
        if not self.period_save:
            return
        tsim = self.sim.time_stepping.t
        if self.due(tsim):
            self.t_last_save = tsim
            self._save(self.compute())

    def _save(self, results):
        self.times.append(self.sim.time_stepping.t)
        self.records.append(results)
        self.nb_saved_times += 1
        if self.output._has_to_save and self.path_file:
            self._write_npz(self.path_file, self.records)

    def _write_npz(self, path, records):
        arrays = {"times": np.array(self.times)}
        arrays.update({k: np.asarray(v) for k, v in self.arrays_1st_time.items()})
        for key in records[-1]:
            arrays[key] = np.stack([np.asarray(r[key]) for r in records])
        tmp = path + ".tmp.npz"
        np.savez(tmp, **arrays)
        os.replace(tmp, path)

    def compute(self):
        raise NotImplementedError

    # loading ---------------------------------------------------------
    def load(self, records=None):
        """Stacked arrays of every saved time (memory, else the file)."""
        records = self.records if records is None else records
        if records:
            out = {"times": np.array(self.times)}
            out.update(self.arrays_1st_time)
            for key in records[-1]:
                out[key] = np.stack([np.asarray(r[key]) for r in records])
            return out
        if self.path_file and os.path.exists(self.path_file):
            with np.load(self.path_file) as f:
                return {k: f[k] for k in f.files}
        raise ValueError(f"no saved {self._tag} data")

    def _mean(self, data, tmin=None, tmax=None):
        t = data["times"]
        sel = np.ones(t.shape, bool)
        if tmin is not None:
            sel &= t >= tmin
        if tmax is not None:
            sel &= t <= tmax
        return {k: (v[sel].mean(axis=0) if isinstance(v, np.ndarray) and v.shape[:1] == t.shape
                    and k != "times" else v) for k, v in data.items()}


class SpectraBase(SpecificOutput):
    _tag = "spectra"

    def __init__(self, output):
        super().__init__(output, output.params.output.periods_save.spectra)


class SpectralEnergyBudgetBase(SpecificOutput):
    _tag = "spect_energy_budg"
    _name_file = "spect_energy_budg.npz"

    def __init__(self, output, arrays_1st_time=None):
        super().__init__(output, output.params.output.periods_save.spect_energy_budg,
                         arrays_1st_time)


class SpatialMeansBase(SpecificOutput):
    """spatial_means.txt in the reference's text format
    (fluidsim/base/output/spatial_means.py; the per-solver
    _save_one_time methods write the blocks)."""

    _tag = "spatial_means"
    _name_file = "spatial_means.txt"

    def __init__(self, output):
        super().__init__(output, output.params.output.periods_save.spatial_means)

    def _save(self, results):
        self.times.append(self.sim.time_stepping.t)
        self.records.append(results)
        self.nb_saved_times += 1
        if self.output._has_to_save:
            with open(self.path_file, "a") as f:
                f.write(self._format(self.sim.time_stepping.t, results))

    def compute(self):
        raise NotImplementedError

    def _format(self, t, r):
        raise NotImplementedError

    def load(self):
        """Dict of arrays keyed as in the file (t, E, ..., *_tot)."""
        if self.records:
            keys = self.records[-1].keys()
            out = {"t": np.array(self.times)}
            out.update({k: np.array([r[k] for r in self.records]) for k in keys})
            return out
        return parse_spatial_means(self.path_file)


def parse_spatial_means(path):
    """Read a spatial_means.txt of '####' blocks of 'name = value' pairs."""
    blocks = []
    with open(path) as f:
        cur = None
        for line in f:
            line = line.strip()
            if line.startswith("####"):
                cur = {}
                blocks.append(cur)
                continue
            if cur is None:
                continue
            for part in line.split(";"):
                if "=" in part:
                    k, v = part.split("=")
                    k = "t" if k.strip() == "time" else k.strip()
                    cur[k] = float(v)
    keys = blocks[0].keys() if blocks else []
    return {k: np.array([b[k] for b in blocks]) for k in keys}


def e(x):
    return f"{x:11.5e}"
