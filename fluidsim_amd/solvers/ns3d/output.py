"""Output of the ns3d solver (fluidsim/solvers/ns3d/output/__init__.py:101-117,
spatial_means.py:23-77, spectra.py:15-99, spect_energy_budget.py:106-190),
computed on the device."""

import numpy as np

from ...base import diagnostics as dg
from ...base.output import OutputBasePseudoSpectral


def spectra_axes(oper):
    """Wavenumber axes of the 1D, 3D and kz-kh spectra."""
    nkx, nky, nkz = oper.nx // 2 + 1, oper.ny // 2 + 1, oper.nz // 2 + 1
    dk = max(oper.deltakx, oper.deltaky, oper.deltakz)
    kmax = min(oper.deltakx * nkx, oper.deltaky * nky, oper.deltakz * nkz)
    dkh = max(oper.deltakx, oper.deltaky)
    khmax = min(oper.deltakx * nkx, oper.deltaky * nky)
    return {
        "kx": oper.deltakx * np.arange(nkx),
        "ky": oper.deltaky * np.arange(nky),
        "kz": oper.deltakz * np.arange(nkz),
        "k_spectra3d": dk * np.arange(max(2, int(kmax / dk))),
        "kh_spectra": dkh * np.arange(max(2, int(khmax / dkh))),
    }


class SpatialMeansNS3D(dg.SpatialMeansBase):
    """SpatialMeansNS3D._save_one_time (spatial_means.py:23-77)."""

    def compute(self):
        sim = self.sim
        g = sim.state.state_spect.get_var
        fields = [g("vx_fft"), g("vy_fft"), g("vz_fft")]
        forcing = getattr(sim, "is_forcing_enabled", False)
        if forcing:
            fg = sim.forcing.get_forcing().get_var
            fields += [fg("vx_fft"), fg("vy_fft"), fg("vz_fft")]
        s = dg.run(self.oper, dg.NS3D_MEANS, 0, fields, dg.diss_params(sim))["sums"]
        p = sim.params
        r = {"E": s[0] + s[1] + s[2], "Ex": s[0], "Ey": s[1], "Ez": s[2],
             "epsK": s[3], "epsK_hypo": s[4], "epsK_tot": s[3] + s[4]}
        if p.nu_4 > 0:
            r["epsK4"] = s[5]
        if p.nu_8 > 0:
            r["epsK8"] = s[6]
        if forcing:
            dt2 = sim.time_stepping.deltat / 2
            r.update({"PK1": s[7], "PK2": dt2 * s[8], "PK_tot": s[7] + dt2 * s[8]})
        return r

    def _format(self, t, r):
        e = dg.e
        txt = (f"####\ntime = {e(t)}\n"
               f"E    = {e(r['E'])}\n"
               f"Ex   = {e(r['Ex'])} ; Ey   = {e(r['Ey'])} ; Ez   = {e(r['Ez'])}\n"
               f"epsK = {e(r['epsK'])} ; epsK_hypo = {e(r['epsK_hypo'])} ; "
               f"epsK_tot = {e(r['epsK_tot'])} \n")
        if "epsK4" in r:
            txt += f"epsK4 = {e(r['epsK4'])}\n"
        if "epsK8" in r:
            txt += f"epsK8 = {e(r['epsK8'])}\n"
        if "PK1" in r:
            txt += (f"PK1  = {e(r['PK1'])} ; PK2       = {e(r['PK2'])} ; "
                    f"PK_tot   = {e(r['PK_tot'])} \n")
        return txt


class SpectraNS3D(dg.SpectraBase):
    """SpectraNS3D.compute (spectra.py:15-99): 1D spectra along kx, ky, kz,
    3D spectra and, every kzkh_periodicity saves, kz-kh spectra."""

    _kind = dg.NS3D_SPECTRA
    _names = ("vx", "vy", "vz", "Khr", "Khd")

    def __init__(self, output):
        super().__init__(output)
        self.kzkh_periodicity = int(self.params.output.spectra.kzkh_periodicity or 0)
        axes = spectra_axes(self.oper)
        self.k_spectra3d = axes["k_spectra3d"]
        self.kh_spectra = axes["kh_spectra"]
        self.arrays_1st_time = {"kx": axes["kx"], "ky": axes["ky"], "kz": axes["kz"],
                                "k_spectra3d": self.k_spectra3d, "kh_spectra": self.kh_spectra}
        self._kzkh_records = []
        self._kzkh_times = []

    def has_to_save_kzkh(self):
        """spectra3d.py:160-171."""
        return bool(self.kzkh_periodicity) and (
            self.nb_saved_times % self.kzkh_periodicity == 0)

    def _fields(self):
        g = self.sim.state.state_spect.get_var
        return [g("vx_fft"), g("vy_fft"), g("vz_fft")]

    def _params(self):
        return dg.diss_params(self.sim)

    def compute(self, kzkh=None):
        if kzkh is None:
            kzkh = self.has_to_save_kzkh()
        flags = dg.KX | dg.KY | dg.KZ | dg.SHELL | (dg.KZKH if kzkh else 0)
        r = dg.run(self.oper, self._kind, flags, self._fields(), self._params())
        idx = {n: i for i, n in enumerate(self._names)}
        d1, d3 = {}, {}
        for n in ("vx", "vy", "vz"):
            for ax in ("kx", "ky", "kz"):
                d1[f"{n}_{ax}"] = r[ax][idx[n]]
        for ax in ("kx", "ky", "kz"):
            d1[f"E_{ax}"] = d1[f"vx_{ax}"] + d1[f"vy_{ax}"] + d1[f"vz_{ax}"]
        for n in self._names[3:]:
            for ax in ("kx", "ky", "kz"):
                d1[f"{n}_{ax}"] = r[ax][idx[n]]
        for n in ("vx", "vy", "vz"):
            d3[n] = r["shell"][idx[n]]
        d3["E"] = d3["vx"] + d3["vy"] + d3["vz"]
        for n in self._names[3:]:
            d3[n] = r["shell"][idx[n]]
        d1 = {"spectra_" + k: v for k, v in d1.items()}
        d3 = {"spectra_" + k: v for k, v in d3.items()}
        dkzkh = self._kzkh(r["kzkh"], idx) if kzkh else None
        return d1, d3, dkzkh

    def _kzkh(self, h, idx):
        return {"K": h[idx["vx"]] + h[idx["vy"]] + h[idx["vz"]], "Khr": h[idx["Khr"]],
                "Khd": h[idx["Khd"]]}

    def _save(self, results):
        d1, d3, dk = results
        t = self.sim.time_stepping.t
        self.times.append(t)
        self.records.append({**d1, **d3})
        if dk is not None:
            self._kzkh_times.append(t)
            self._kzkh_records.append(dk)
        self.nb_saved_times += 1
        if self.output._has_to_save:
            path = self.output.path_run
            self._write_npz(path + "/spectra1d.npz",
                            [{k: r[k] for k in d1} for r in self.records])
            self._write_npz(path + "/spectra3d.npz",
                            [{k: r[k] for k in d3} for r in self.records])
            if dk is not None:
                arrays = {"times": np.array(self._kzkh_times), **self.arrays_1st_time}
                for k in dk:
                    arrays[k] = np.stack([r[k] for r in self._kzkh_records])
                np.savez(path + "/spectra_kzkh.npz", **arrays)

    def load1d_mean(self, tmin=None, tmax=None):
        d = self._mean(self.load(), tmin, tmax)
        return {k: v for k, v in d.items() if not k.startswith("spectra_") or
                k.endswith(("_kx", "_ky", "_kz"))}

    def load3d_mean(self, tmin=None, tmax=None):
        d = self._mean(self.load(), tmin, tmax)
        return {k: v for k, v in d.items() if not k.endswith(("_kx", "_ky", "_kz"))}

    def load_kzkh_mean(self, tmin=None, tmax=None):
        if not self._kzkh_records:
            raise ValueError("no saved kz-kh spectra")
        data = {"times": np.array(self._kzkh_times)}
        for k in self._kzkh_records[-1]:
            data[k] = np.stack([r[k] for r in self._kzkh_records])
        return self._mean(data, tmin, tmax)


class SpectralEnergyBudgetNS3D(dg.SpectralEnergyBudgetBase):
    """SpectralEnergyBudgetNS3D.compute (spect_energy_budget.py:106-190):
    transfer_Kh, transfer_Kz, transfer_Khr, diss_Kh, diss_Kz as kz-kh
    spectra and 1D spectra (_kx, _ky, _kz)."""

    _kind = dg.NS3D_BUDGET
    _names = ("transfer_Kh", "transfer_Kz", "transfer_Khr", "diss_Kh", "diss_Kz")

    def __init__(self, output):
        axes = spectra_axes(output.sim.oper)
        super().__init__(output, {"kx": axes["kx"], "ky": axes["ky"], "kz": axes["kz"],
                                  "kh": axes["kh_spectra"]})

    def _transfer_fields(self):
        """P fft(v x omega) without Coriolis (the budget's own product,
        spect_energy_budget.py:128-152), projected by the solver's projector."""
        import torch

        sim, oper = self.sim, self.oper
        S = sim.state.state_spect.data
        f = torch.zeros((3,) + tuple(S.shape[1:]), dtype=S.dtype, device=S.device)
        oper.ns3d_tendency(S[:3], f, f=None)
        proj = {1: oper.project_perpk3d, 2: oper.project_toroidal,
                3: oper.project_poloidal}[sim._engine_projection]
        proj(f[0], f[1], f[2])
        return f

    def _fields(self):
        S = self.sim.state.state_spect.data
        f = self._transfer_fields()
        return [S[0], S[1], S[2], f[0], f[1], f[2]], f

    def _params(self):
        return dg.diss_params(self.sim)

    def compute(self):
        fields, _keep = self._fields()
        flags = dg.KX | dg.KY | dg.KZ | dg.KZKH
        r = dg.run(self.oper, self._kind, flags, fields, self._params())
        res = {}
        for i, n in enumerate(self._names):
            res[n] = r["kzkh"][i]
            res[n + "_kx"], res[n + "_ky"], res[n + "_kz"] = (r["kx"][i], r["ky"][i],
                                                            r["kz"][i])
        return res


class OutputNS3D(OutputBasePseudoSpectral):
    _cls_spatial_means = SpatialMeansNS3D
    _cls_spectra = SpectraNS3D
    _cls_spect_energy_budg = SpectralEnergyBudgetNS3D

    @staticmethod
    def _complete_params_with_default(params, info_solver=None):
        OutputBasePseudoSpectral._complete_params_with_default(params, info_solver)
        params.output.phys_fields.field_to_plot = "rotz"
        params.output.spectra._set_attrib("kzkh_periodicity", 0)

    def _v(self):
        g = self.sim.state.state_spect.get_var
        return [g("vx_fft"), g("vy_fft"), g("vz_fft")]

    def compute_energy(self):
        return self._sums(1, self._v())[0]

    def compute_energies(self):
        """(E_vx, E_vy, E_vz): Parseval sums of compute_energies_fft
        (ns3d/output/__init__.py:101-109)."""
        s = dg.run(self.oper, dg.NS3D_MEANS, 0, self._v(), dg.diss_params(self.sim))["sums"]
        return s[0], s[1], s[2]

    def compute_energy_spectrum(self):
        """(k_spectra3d, 3D spectrum of the kinetic energy)."""
        return self.spectra.k_spectra3d, self.spectra.compute(kzkh=False)[1]["spectra_E"]
