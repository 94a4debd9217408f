"""Output of the ns3d.strat solver: kinetic and potential energies
E_K, E_A = |b|^2 / (2 N^2) and their dissipation (FP64 sums)."""

from ..output import OutputNS3D


class OutputNS3DStrat(OutputNS3D):
    def _b(self):
        return self.sim.state.state_spect.get_var("b_fft")

    def compute_energies(self):
        EK, EA = self._sums(2, self._v() + [self._b()], self.sim.params.N ** 2)[:2]
        return EK, EA

    def compute_energy(self):
        EK, EA = self.compute_energies()
        return EK + EA

    def compute_spatial_means(self):
        EK, EA, epsK, epsA = self._sums(2, self._v() + [self._b()],
                                        self.sim.params.N ** 2)[:4]
        return {"E": EK + EA, "EK": EK, "EA": EA, "epsK": epsK, "epsA": epsA,
                "eps_tot": epsK + epsA}
