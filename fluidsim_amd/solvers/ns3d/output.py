"""Output of the ns3d solver (fluidsim/solvers/ns3d/output/__init__.py:101-117,
spatial_means.py:23-73 energy and dissipation)."""

from ...base.output import OutputBasePseudoSpectral


class OutputNS3D(OutputBasePseudoSpectral):
    @staticmethod
    def _complete_params_with_default(params, info_solver=None):
        OutputBasePseudoSpectral._complete_params_with_default(params, info_solver)
        params.output.phys_fields.field_to_plot = "rotz"

    def _v(self):
        g = self.sim.state.state_spect.get_var
        return [g("vx_fft"), g("vy_fft"), g("vz_fft")]

    def compute_energy(self):
        return self._sums(1, self._v())[0]

    def compute_spatial_means(self):
        E, epsK, epsK_hypo = self._sums(1, self._v())[:3]
        return {"E": E, "epsK": epsK, "epsK_hypo": epsK_hypo,
                "epsK_tot": epsK + epsK_hypo}

    def compute_energy_spectrum(self, nbins=None):
        return self.compute_spectrum(self._v(), nbins)
