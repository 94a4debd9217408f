"""fluidsim_amd: MI355X-native engine for fluidsim's ns2d / ns3d pseudo-spectral
time step.

Drop-in usage (same API as fluidsim)::

    from fluidsim_amd.solvers.ns2d.solver import Simul
    params = Simul.create_default_params()
    ...
    sim = Simul(params)
    sim.time_stepping.start()

State arrays live in HBM as torch tensors; the hot path runs in hand-written
gfx950 HIP kernels reached through the C-ABI in include/fluidsim_amd.h.
"""

__version__ = "0.1.0"
