"""Output of the ns2d solver (fluidsim/solvers/ns2d/output/__init__.py:110-129,
spatial_means.py:23-68, spectra.py:16-28, spect_energy_budget.py:23-80),
computed on the device."""

import numpy as np

from ...base import diagnostics as dg
from ...base.output import OutputBasePseudoSpectral


class SpatialMeansNS2D(dg.SpatialMeansBase):
    """SpatialMeansNS2D._save_one_time (spatial_means.py:23-68)."""

    def compute(self):
        sim = self.sim
        rot = sim.state.state_spect.get_var("rot_fft")
        fields = [rot]
        forcing = getattr(sim, "is_forcing_enabled", False)
        if forcing:
            fields.append(sim.forcing.get_forcing().get_var("rot_fft"))
        s = dg.run(self.oper, dg.NS2D_MEANS, 0, fields, dg.diss_params(sim))["sums"]
        r = {"E": s[0], "Z": s[1], "epsK": s[2], "epsK_hypo": s[3],
             "epsK_tot": s[2] + s[3], "epsZ": s[4], "epsZ_hypo": s[5],
             "epsZ_tot": s[4] + s[5]}
        if forcing:
            dt2 = sim.time_stepping.deltat / 2
            r.update({"PK1": s[6], "PK2": dt2 * s[7], "PK_tot": s[6] + dt2 * s[7],
                      "PZ1": s[8], "PZ2": dt2 * s[9], "PZ_tot": s[8] + dt2 * s[9]})
        return r

    def _format(self, t, r):
        e = dg.e
        txt = (f"####\ntime = {e(t)}\n"
               f"E    = {e(r['E'])} ; Z         = {e(r['Z'])} \n"
               f"epsK = {e(r['epsK'])} ; epsK_hypo = {e(r['epsK_hypo'])} ; "
               f"epsK_tot = {e(r['epsK_tot'])} \n"
               f"epsZ = {e(r['epsZ'])} ; epsZ_hypo = {e(r['epsZ_hypo'])} ; "
               f"epsZ_tot = {e(r['epsZ_tot'])} \n")
        if "PK1" in r:
            txt += (f"PK1  = {e(r['PK1'])} ; PK2       = {e(r['PK2'])} ; "
                    f"PK_tot   = {e(r['PK_tot'])} \n"
                    f"PZ1  = {e(r['PZ1'])} ; PZ2       = {e(r['PZ2'])} ; "
                    f"PZ_tot   = {e(r['PZ_tot'])} \n")
        return txt


class SpectraNS2D(dg.SpectraBase):
    """SpectraNS2D.compute (spectra.py:16-28): 1D spectra along kx and ky
    and the 2D (shell) spectrum of the kinetic energy."""

    def __init__(self, output):
        super().__init__(output)
        oper = self.oper
        self.kxE = oper.deltakx * np.arange(oper.nx // 2 + 1)
        self.kyE = oper.deltaky * np.arange(oper.ny // 2 + 1)
        self.khE = self.kxE
        self.arrays_1st_time = {"kxE": self.kxE, "kyE": self.kyE, "khE": self.khE}

    def compute(self):
        rot = self.sim.state.state_spect.get_var("rot_fft")
        r = dg.run(self.oper, dg.NS2D_SPECTRA, dg.KX | dg.KY | dg.SHELL, [rot])
        return ({"spectrum1Dkx_E": r["kx"][0], "spectrum1Dky_E": r["ky"][0]},
                {"spectrum2D_E": r["shell"][0]})

    def _save(self, results):
        d1, d2 = results
        self.times.append(self.sim.time_stepping.t)
        self.records.append({**d1, **d2})
        self.nb_saved_times += 1
        if self.output._has_to_save:
            path = self.output.path_run
            self._write_npz(path + "/spectra1D.npz",
                            [{k: r[k] for k in d1} for r in self.records])
            self._write_npz(path + "/spectra2D.npz",
                            [{k: r[k] for k in d2} for r in self.records])

    def load1d_mean(self, tmin=None, tmax=None):
        return self._mean(self.load(), tmin, tmax)

    load2d_mean = load1d_mean


class SpectralEnergyBudgetNS2D(dg.SpectralEnergyBudgetBase):
    """SpectralEnergyBudgetNS2D.compute (spect_energy_budget.py:23-80):
    transfer2D_E, transfer2D_Z (shell spectra of the transfer terms)."""

    def __init__(self, output):
        oper = output.sim.oper
        super().__init__(output, {"khE": oper.deltakx * np.arange(oper.nx // 2 + 1)})

    def compute(self):
        import torch

        sim, oper = self.sim, self.oper
        rot = sim.state.state_spect.get_var("rot_fft")
        ct = rot.dtype
        rt = torch.float32 if ct == torch.complex64 else torch.float64
        wk = torch.empty((8,) + tuple(rot.shape), dtype=ct, device=rot.device)
        wx = torch.empty((8, oper.ny, oper.nx), dtype=rt, device=rot.device)
        F = torch.zeros((3,) + tuple(rot.shape), dtype=ct, device=rot.device)
        beta = float(getattr(sim.params, "beta", 0.0) or 0.0)
        dg._lib.call("fsx_ns2d_budget_fields", oper.plan.handle, dg._lib.dptr(rot), beta,
                     dg._lib.dptr(wk), dg._lib.dptr(wx), dg._lib.dptr(F),
                     dg._lib.stream_ptr())
        del wk, wx
        r = dg.run(oper, dg.NS2D_BUDGET, dg.SHELL, [rot, F[0], F[1], F[2]])
        return {"transfer2D_E": r["shell"][0], "transfer2D_Z": r["shell"][1]}


class OutputNS2D(OutputBasePseudoSpectral):
    _cls_spatial_means = SpatialMeansNS2D
    _cls_spectra = SpectraNS2D
    _cls_spect_energy_budg = SpectralEnergyBudgetNS2D

    @staticmethod
    def _complete_params_with_default(params, info_solver=None):
        OutputBasePseudoSpectral._complete_params_with_default(params, info_solver)
        params.output.phys_fields.field_to_plot = "rot"

    def _rot(self):
        return self.sim.state.state_spect.get_var("rot_fft")

    def compute_energy(self):
        return self._sums(0, [self._rot()])[0]

    def compute_enstrophy(self):
        return self._sums(0, [self._rot()])[1]

    def compute_energy_spectrum(self):
        """(khE, 2D spectrum of the kinetic energy)."""
        return self.spectra.khE, self.spectra.compute()[1]["spectrum2D_E"]
