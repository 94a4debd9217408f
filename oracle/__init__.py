"""FP64 CPU oracle for the ns2d / ns3d / ns3d.strat pseudo-spectral time step.

TEST INFRASTRUCTURE ONLY. This package is a plain numpy/scipy restatement of
fluidsim's algorithm for this path (each function cites the reference file:line
it follows). Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline -- never as a fallback for the GPU engine.

Pinning (what fixes each function without importing fluidsim, whose
dependencies are absent here): see the pin column below and the tests in
``tests/test_oracle_*.py``.

==========================  ===================================================
function                    pinned by
==========================  ===================================================
Oper2D/Oper3D.fft/ifft      brute-force DFT at n<=12, Parseval, round trip
sum_wavenumbers             Parseval identity (fluidsim/base/state.py:385-392)
curl / vecfft / grad        round trip rtol 1e-15 (test_operators2d.py:69-80),
                            single-mode closed forms
project_perpk3d/perp        k.P(v)=0, idempotence, single-mode closed form
project_toroidal/poloidal   radial+poloidal+toroidal = id (test_operators3d.py
                            :353-480)
dealiasing masks            index predicate, brute force on small grids
ns2d tendency               energy/enstrophy conservation (ns2d/test_solver.py
                            :56-69, 181-193), Taylor-Green N=0, single-mode
                            Jacobian closed form, planted-mistake tests
ns3d tendency               energy conservation (ns3d/test_solver.py:80-96),
                            ABC flow N=0, closed-form two-mode product
strat tendency              energy incl. E_A conservation (strat/test_solver.py
                            :35-58), linear wave closed form
IF-RK4 / RK2                N=0 exact decay, 4th/2nd order self-convergence,
                            linear ODE closed form per stage
CFL                         scripted maxima, hysteresis rule
energy budget               dE/dt = -eps to O(dt^4)
spectra binning             restated from fluidfft's published algorithm; full-
                            grid brute force, sum rules, single-mode share
                            (parity vs fluidfft itself unpinned: not in tree)
budgets / spatial means     rate of change of the spectra under the oracle's
                            own step (order 1), injection = forcing rate
Euler / RK2 / phase shifts  exact IF decay, order of convergence, alias-free
                            reduction of the shifted tendency
forcing (tcrandom, aniso.)  injection = forcing rate (closed form), ring
                            count, root choice vs numpy.roots
Smagorinsky                 energy identity, traceless strain, shear closed
                            form, brute-force DFT on 4^3
cnn (closure network)       FFT correlation form of the periodic convolution,
                            finite-difference gradients of every weight, bias
                            and the input, unflipped-kernel mistake fails;
                            rotation / mirror symmetry of the ns2d tendency
                            (the training augmentation)
==========================  ===================================================
"""

from .operators import Oper2D, Oper3D  # noqa: F401
