"""Output of the ns3d.strat solver (fluidsim/solvers/ns3d/strat/output/
__init__.py:55-92, spatial_means.py:27-106, spectra.py:18-112,
spect_energy_budget.py:60-118), computed on the device."""

from ....base import diagnostics as dg
from ..output import OutputNS3D, SpectraNS3D, SpectralEnergyBudgetNS3D


class SpatialMeansNS3DStrat(dg.SpatialMeansBase):
    """SpatialMeansNS3DStrat._save_one_time (spatial_means.py:27-106)."""

    def compute(self):
        sim = self.sim
        g = sim.state.state_spect.get_var
        fields = [g("vx_fft"), g("vy_fft"), g("vz_fft"), g("b_fft")]
        forcing = getattr(sim, "is_forcing_enabled", False)
        if forcing:
            fg = sim.forcing.get_forcing().get_var
            fields += [fg("vx_fft"), fg("vy_fft"), fg("vz_fft"), fg("b_fft")]
        N = sim.params.N
        s = dg.run(self.oper, dg.STRAT_MEANS, 0, fields, dg.diss_params(sim, N))["sums"]
        p = sim.params
        r = {"EA": s[0], "EKz": s[1], "EKhr": s[2] - s[4], "EKhd": s[3], "EKhs": s[4],
             "EAs": s[5]}
        r["E"] = r["EA"] + r["EKz"] + r["EKhr"] + r["EKhd"] + r["EKhs"]
        r.update({"epsK": s[6], "epsK_hypo": s[7], "epsA": s[8], "epsA_hypo": s[9]})
        r["eps_tot"] = s[6] + s[7] + s[8] + s[9]
        if p.nu_4 > 0:
            r["epsK4"], r["epsA4"] = s[10], s[11]
        if p.nu_8 > 0:
            r["epsK8"], r["epsA8"] = s[12], s[13]
        if forcing:
            dt2 = sim.time_stepping.deltat / 2
            r.update({"PK1": s[14], "PK2": dt2 * s[15], "PK_tot": s[14] + dt2 * s[15],
                      "PA1": s[16] / N**2, "PA2": dt2 * s[17] / N**2,
                      "PA_tot": (s[16] + dt2 * s[17]) / N**2})
        return r

    def _format(self, t, r):
        e = dg.e
        txt = (f"####\ntime = {e(t)}\n"
               f"E    = {e(r['E'])}\n"
               f"EA   = {e(r['EA'])} ; EKz   = {e(r['EKz'])} ; "
               f"EKhr   = {e(r['EKhr'])} ; EKhd   = {e(r['EKhd'])} ; "
               f"EKhs   = {e(r['EKhs'])} ; EAs    = {e(r['EAs'])}\n"
               f"epsK = {e(r['epsK'])} ; epsK_hypo = {e(r['epsK_hypo'])} ; "
               f"epsA = {e(r['epsA'])} ; epsA_hypo = {e(r['epsA_hypo'])} ; "
               f"eps_tot = {e(r['eps_tot'])} \n")
        if "epsK4" in r:
            txt += f"epsK4 = {e(r['epsK4'])} ; epsA4 = {e(r['epsA4'])}\n"
        if "epsK8" in r:
            txt += f"epsK8 = {e(r['epsK8'])} ; epsA8 = {e(r['epsA8'])}\n"
        if "PK1" in r:
            txt += (f"PK1  = {e(r['PK1'])} ; PK2       = {e(r['PK2'])} ; "
                    f"PK_tot   = {e(r['PK_tot'])} \n"
                    f"PA1  = {e(r['PA1'])} ; PA2       = {e(r['PA2'])} ; "
                    f"PA_tot   = {e(r['PA_tot'])} \n")
        return txt


class SpectraNS3DStrat(SpectraNS3D):
    """SpectraNS3DStrat.compute (spectra.py:18-112): adds the potential
    energy A = |b|^2 / (2 N^2)."""

    _kind = dg.STRAT_SPECTRA
    _names = ("vx", "vy", "vz", "A", "Khr", "Khd")

    def _fields(self):
        g = self.sim.state.state_spect.get_var
        return [g("vx_fft"), g("vy_fft"), g("vz_fft"), g("b_fft")]

    def _params(self):
        return dg.diss_params(self.sim, self.sim.params.N)

    def _kzkh(self, h, idx):
        return {"A": h[idx["A"]], "Khr": h[idx["Khr"]], "Khd": h[idx["Khd"]],
                "Kz": h[idx["vz"]]}


class SpectralEnergyBudgetNS3DStrat(SpectralEnergyBudgetNS3D):
    """SpectralEnergyBudgetNS3DStrat.compute (spect_energy_budget.py:60-118):
    adds diss_A, conv_K2A and transfer_A (from the b tendency of the
    engine, -i k.(v b)^ - N^2 vz)."""

    _kind = dg.STRAT_BUDGET
    _names = SpectralEnergyBudgetNS3D._names + ("diss_A", "conv_K2A", "transfer_A")

    def _fields(self):
        import torch

        sim, oper = self.sim, self.oper
        S = sim.state.state_spect.data
        f = self._transfer_fields()
        t = torch.zeros_like(S)
        dg._lib.call("fsx_strat_tendency", oper.plan.handle, dg._lib.dptr(S),
                     dg._lib.dptr(t), 0, 0.0, float(sim.params.N), 0, 0, 0,
                     dg._lib.stream_ptr())
        return [S[0], S[1], S[2], f[0], f[1], f[2], S[3], t[3]], (f, t)

    def _params(self):
        return dg.diss_params(self.sim, self.sim.params.N)


class OutputNS3DStrat(OutputNS3D):
    _cls_spatial_means = SpatialMeansNS3DStrat
    _cls_spectra = SpectraNS3DStrat
    _cls_spect_energy_budg = SpectralEnergyBudgetNS3DStrat

    def _b(self):
        return self.sim.state.state_spect.get_var("b_fft")

    def compute_energies(self):
        """(E_A, E_Kz, E_Khr, E_Khd) (strat/output/__init__.py:55-72, 86-88)."""
        s = dg.run(self.oper, dg.STRAT_MEANS, 0, self._v() + [self._b()],
                   dg.diss_params(self.sim, self.sim.params.N))["sums"]
        return s[0], s[1], s[2], s[3]

    def compute_energy(self):
        EK, EA = self._sums(2, self._v() + [self._b()], self.sim.params.N ** 2)[:2]
        return EK + EA
