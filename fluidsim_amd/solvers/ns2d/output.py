"""Output of the ns2d solver (fluidsim/solvers/ns2d/output/__init__.py:110-129,
spatial_means.py:23-95 without forcing)."""

from ...base.output import OutputBasePseudoSpectral


class OutputNS2D(OutputBasePseudoSpectral):
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

    def compute_spatial_means(self):
        E, Z, epsK, epsK_hypo, epsZ, epsZ_hypo = self._sums(0, [self._rot()])
        return {
            "E": E, "Z": Z, "epsK": epsK, "epsK_hypo": epsK_hypo,
            "epsK_tot": epsK + epsK_hypo, "epsZ": epsZ, "epsZ_hypo": epsZ_hypo,
            "epsZ_tot": epsZ + epsZ_hypo,
        }

    def compute_energy_spectrum(self, nbins=None):
        """Shell spectrum of the kinetic energy (ux, uy from rot)."""
        ux_fft, uy_fft = self.oper.vecfft_from_rotfft(self._rot())
        return self.compute_spectrum([ux_fft.contiguous(), uy_fft.contiguous()], nbins)
