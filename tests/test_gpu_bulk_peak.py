"""GPU parity of the bulk viscosity with zeta/s peaked at T_c
(echo_set_bulk_peak; the oracle's model is pinned in
tests/test_oracle_bulk_peak.py), through the C-ABI, against the oracle on
the same seeded inputs: viscous hydro (RK3) and viscous resistive MHD
(IMEX), and the Navier-Stokes bulk pressure of echo_init_glauber.  Bar:
relative L-infinity <= 1e-10 in FP64 (1e-12 for the initial state, whose
only difference is the last bit of exp)."""
import numpy as np
import pytest

from libs import make_config, oracle_glauber_prim
from scenarios import fireball3d, rel_linf, rmhd_fireball

pytestmark = pytest.mark.gpu
TOL = 1e-10
ZS, TC, W = 0.08, 0.155, 0.02
PAR = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=5.36, deta=3.0, sigeta=0.8, rhocenter=1.0)

CASES = {
    "viscous-rk3": lambda: fireball3d(24, 20, 16, viscosity=1, eqstate=3, bulk=1),
    "vrmhd-imex": lambda: rmhd_fireball(viscosity=1, nrk=4, eqstate=3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_bulk_peak_evolution(oracle, name):
    from echoamd import EchoGrid
    ic, rc, v0 = CASES[name]()
    assert ic[5] == 1   # BULK
    oracle.setup(ic, rc)
    oracle.set_bulk_peak(ZS, TC, W)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(5) == 0
    g = EchoGrid(ic, rc)
    g.set_bulk_peak(ZS, TC, W)
    assert g.init_from_prim(v0) == 0
    g.steps(5)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL
    assert np.abs(vo[5]).max() > 0
    g.close()


def test_bulk_peak_glauber_initial_state(oracle):
    from echoamd import EchoGrid
    rng = np.random.default_rng(12)
    ic, rc = make_config(nx=40, ny=36, nz=16, coordinates=2, viscosity=1, bulk=1, eqstate=3,
                         xlim=(-6.0, 6.0), ylim=(-5.4, 5.4), zlim=(-4.0, 4.0), tmin=0.6,
                         init_type=3, eta_over_s=0.12, maxdt=0.05)
    pa = rng.normal(0, 2.0, (150, 2)) + [1.0, 0.0]
    pb = rng.normal(0, 2.0, (140, 2)) - [1.0, 0.0]
    pc = rng.normal(0, 1.6, (420, 2))
    oracle.setup(ic, rc)
    oracle.set_bulk_peak(ZS, TC, W)
    e, v0 = oracle_glauber_prim(oracle, PAR, pa, pb, pc)
    assert e == 0 and np.abs(v0[5]).max() > 0
    assert oracle.init_from_prim(v0) == 0
    vo, uo = oracle.state()
    g = EchoGrid(ic, rc)
    g.set_bulk_peak(ZS, TC, W)
    assert g.init_glauber(PAR, pa, pb, pc) == 0
    vg, ug = g.state()
    assert rel_linf(vg, vo) <= 1e-12 and rel_linf(ug, uo) <= 1e-12
    assert oracle.steps(3) == 0
    g.steps(3)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL
    g.close()


def test_bulk_peak_rejected_without_bulk():
    from echoamd import EchoError, EchoGrid
    ic, rc, v0 = fireball3d(12, 10, 8, viscosity=1, eqstate=3, bulk=0)
    g = EchoGrid(ic, rc)
    with pytest.raises(EchoError, match="BULK"):
        g.set_bulk_peak(ZS, TC, W)
    g.close()
