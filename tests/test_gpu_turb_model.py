"""GPU parity of the Smagorinsky closure hook
(fluidsim/base/turb_model/smagorinsky.py:18-110; ns3d/solver.py:246-249)
against the FP64 oracle.  Bars: model term 1e-12 relative L2 and exact
energy identity (Parseval sum of Re(conj(v) f) = -4 C_nuT mean(|S|^3)) to
1e-11; FP64 state after CFL steps 1e-11."""

import numpy as np
import pytest

from oracle.operators import Oper3D
from oracle.solvers import NS3D, NS3DStrat
from oracle.turb_model import Smagorinsky
from fluidsim_amd.synthetic import ns3d_spectrum_vv_fft

from .gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def smag_sim(n, strat=False, it_end=4, C=0.18):
    from fluidsim_amd.base.turb_model import SmagorinskyModel, extend_simul_class

    if strat:
        from fluidsim_amd.solvers.ns3d.strat.solver import Simul as S0
    else:
        from fluidsim_amd.solvers.ns3d.solver import Simul as S0
    Simul = extend_simul_class(S0, SmagorinskyModel)
    params = Simul.create_default_params()
    params.oper.nx = params.oper.ny = params.oper.nz = n
    params.oper.precision = "float64"
    params.nu_2 = 1e-3
    if strat:
        params.N = 3.0
    params.turb_model.enable = True
    params.turb_model.type = "smagorinsky"
    params.turb_model.smagorinsky.C = C
    params.time_stepping.USE_CFL = True
    params.time_stepping.cfl_coef = 0.5
    params.time_stepping.USE_T_END = False
    params.time_stepping.it_end = it_end
    params.init_fields.type = "in_script"
    params.output.HAS_TO_SAVE = False
    params.output.periods_print.print_stdout = 0
    return Simul(params)


def test_smagorinsky_term_and_energy_identity():
    n = 32
    sim = smag_sim(n)
    assert "smagorinsky" in sim.info_solver.classes.TurbModel.classes._tag_children
    vv = ns3d_spectrum_vv_fft(n, seed=11)
    sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
    v = sim.state.state_spect
    f = sim.turb_model.get_forcing(vx_fft=v.get_var("vx_fft"), vy_fft=v.get_var("vy_fft"),
                                   vz_fft=v.get_var("vz_fft")).to_numpy()
    o = Oper3D(n, n, n)
    m = Smagorinsky(o, C=0.18)
    ref = m.get_forcing(*vv)
    err = rel_l2(f, np.array(ref))
    S = m.stress_tensor(*vv)
    expected = -4 * m.C_nu_T * np.mean(m.norm(*S) ** 3)
    rate = sum(o.sum_wavenumbers((a.conj() * b).real) for a, b in zip(vv, f))
    print(f"smagorinsky term rel L2 = {err:.3e}; energy rate {rate:.6e} vs {expected:.6e}")
    assert err < 1e-12
    assert rate == pytest.approx(expected, rel=1e-11)


@pytest.mark.parametrize("strat", [False, True])
def test_smagorinsky_steps_vs_oracle(strat):
    n, nsteps = 32, 5
    print(f"case smagorinsky strat={strat} n={n}", flush=True)
    sim = smag_sim(n, strat=strat, it_end=nsteps)
    vv = ns3d_spectrum_vv_fft(n, seed=12)
    if strat:
        b = ns3d_spectrum_vv_fft(n, seed=13)[0] * 0.3
        sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2], b_fft=b)
        orc = NS3DStrat(Oper3D(n, n, n), N=3.0, nu_2=1e-3, USE_CFL=True, cfl_coef=0.5)
        orc.state_spect[:] = np.array(list(vv) + [b])
    else:
        sim.state.init_statespect_from(vx_fft=vv[0], vy_fft=vv[1], vz_fft=vv[2])
        orc = NS3D(Oper3D(n, n, n), nu_2=1e-3, USE_CFL=True, cfl_coef=0.5)
        orc.state_spect[:] = np.array(vv)
    if not strat:
        # ns3d.strat's tendencies have no model term
        # (fluidsim/solvers/ns3d/strat/solver.py:138-216)
        orc.turb_model = Smagorinsky(orc.oper, C=0.18)
    sim.time_stepping.start()
    got = sim.state.state_spect.to_numpy()
    orc.run(nsteps)
    err = rel_l2(got, orc.state_spect)
    print(f"smagorinsky strat={strat}: rel L2 = {err:.3e}")
    assert sim.time_stepping.t == pytest.approx(orc.t, rel=1e-13)
    assert err <= 1e-11
