"""The GPU Glauber-MC initial state (echo_init_glauber) against the oracle
(oracle_init.c followed by the oracle's initialisation), and the first
steps evolved from it."""
import numpy as np
import pytest

from libs import make_config, oracle_glauber_prim
from scenarios import rel_linf

pytestmark = pytest.mark.gpu
PAR = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=5.36, deta=3.0, sigeta=0.8, rhocenter=1.0)


@pytest.mark.parametrize("viscosity,eqstate", [(1, 3), (0, 1), (0, 3)])
def test_glauber_initial_state(oracle, viscosity, eqstate):
    from echoamd import EchoGrid
    rng = np.random.default_rng(11)
    ic, rc = make_config(nx=40, ny=36, nz=16, coordinates=2, viscosity=viscosity, eqstate=eqstate,
                         xlim=(-6.0, 6.0), ylim=(-5.4, 5.4), zlim=(-4.0, 4.0), tmin=0.6,
                         init_type=3, eta_over_s=0.1, maxdt=0.05)
    pa = rng.normal(0, 2.0, (150, 2)) + [1.0, 0.0]
    pb = rng.normal(0, 2.0, (140, 2)) - [1.0, 0.0]
    pc = rng.normal(0, 1.6, (420, 2))
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, PAR, pa, pb, pc)
    assert e == 0
    assert oracle.init_from_prim(v0) == 0
    vo, uo = oracle.state()
    g = EchoGrid(ic, rc)
    assert g.init_glauber(PAR, pa, pb, pc) == 0
    vg, ug = g.state()
    # device exp and glibc exp differ in the last bit; the sums of 710
    # Gaussians keep that at the 1e-15 level
    assert rel_linf(vg, vo) <= 1e-12
    assert rel_linf(ug, uo) <= 1e-12
    assert oracle.steps(3) == 0
    g.steps(3)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= 1e-10 and rel_linf(ug, uo) <= 1e-10
    g.close()


def test_glauber_rejects_bad_input():
    from echoamd import EchoError, EchoGrid
    ic, rc = make_config(nx=8, ny=8, nz=4, coordinates=2)
    g = EchoGrid(ic, rc)
    with pytest.raises(EchoError):
        g.init_glauber(dict(PAR, sig_mc=0.0), [[0, 0]], [[0, 0]], [[0, 0]])
    with pytest.raises(EchoError):
        g.init_glauber(PAR, np.zeros((3000, 2)), np.zeros((1000, 2)), np.zeros((200, 2)))
    g.close()
