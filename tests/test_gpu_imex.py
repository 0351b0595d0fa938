"""IMEX-SSP3(4,3,3) on ideal-MHD grids (INT_ALGO SSP with MHD = 1), which the
reference runs with a zero stiff term (common.f90:2076-2083, evolve.f90:
204-264), through the C-ABI against the oracle, whose IMEX stages equal the
reference's bit for bit on these systems (tests/test_oracle_rmhd.py):

- trajectories: relative L-infinity 1e-10 after four steps, the clock to
  rounding;
- host-forced failures: the repair at every stage with U* of the stage
  start, the oracle's repair count, the states within ten times the oracle's
  own spread under 1e-16 perturbations of its input;
- the stage-by-stage API, the timed steps and an ensemble give the captured
  step's result bit for bit;
- grids without MHD are refused, as is the learned inversion."""
import numpy as np
import pytest

from scenarios import force_superluminal, mhd2d, mhd3d, mhd_fireball3d, rel_linf

pytestmark = pytest.mark.gpu
TOL = 1e-10

CASES = {
    "mhd2d-glm": lambda: mhd2d(24, divclean=1),
    "mhd2d": lambda: mhd2d(24, divclean=0),
    "mhd3d-brentm": lambda: mhd3d(10, 8, 6, algo=1),
    "mhd-fireball-bjorken": lambda: mhd_fireball3d(),
    "mhd3d-e3p": lambda: mhd3d(12, 10, 8, algo=5, eqstate=1),
}


def _imex(t):
    ic, rc, v0 = t
    ic = ic.copy()
    ic[10] = 4
    return ic, rc, v0


def _grid(ic, rc):
    from echoamd import EchoGrid
    return EchoGrid(ic, rc)


@pytest.mark.parametrize("name", sorted(CASES))
def test_imex_ideal_mhd_vs_oracle(oracle, name):
    ic, rc, v0 = _imex(CASES[name]())
    print(f"\nimex {name}", flush=True)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(4) == 0
    vo, uo = oracle.state()
    so = oracle.scalars()
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.steps(4)
    vg, ug = g.state()
    sg = g.scalars()
    assert sg["iter"] == 4
    assert abs(sg["t"] - so["t"]) <= 1e-13 * abs(so["t"])
    assert abs(sg["glm_ch_old"] - so["glm_ch_old"]) <= 1e-10 * abs(so["glm_ch_old"])
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL, (rel_linf(vg, vo), rel_linf(ug, uo))
    g.close()


def _oracle_forced(oracle, ic, rc, v0, pert=0.0, seed=0):
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    v, u = oracle.state()
    u = force_superluminal(u)
    if pert:
        u = u * (1 + pert * np.random.default_rng(seed).standard_normal(u.shape))
    oracle.set_state(v, u)
    assert oracle.steps(2) == 0
    vo, uo = oracle.state()
    return v, u, vo, uo, oracle.repair_counts()["u2v"]


def test_imex_forced_failures(oracle):
    """Superluminal conserved states at three neighbouring cells: every stage
    inverts from the forced u0 (K2 = u0, K3, K4 and the result all contain
    it), so the repair runs at each stage, reading its neighbours' U* through
    the stage start."""
    ic, rc, v0 = _imex(CASES["mhd3d-e3p"]())
    v, u, vo, uo, n = _oracle_forced(oracle, ic, rc, v0)
    assert n >= 3
    spread = 0.0
    for seed in (1, 2, 3):
        _, _, vq, uq, nq = _oracle_forced(oracle, ic, rc, v0, pert=1e-16, seed=seed)
        assert nq == n
        spread = max(spread, rel_linf(vq, vo), rel_linf(uq, uo))
    print(f"\nrepairs {n}, oracle spread under 1e-16 perturbations: {spread:.3e}", flush=True)
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.upload(v, u)
    g.steps(2)
    vg, ug = g.state()
    assert g.repairs() == n
    bar = max(10 * spread, TOL)
    assert rel_linf(vg, vo) <= bar and rel_linf(ug, uo) <= bar, (rel_linf(vg, vo), rel_linf(ug, uo), bar)
    g.close()


def test_imex_stages_timed_and_ensemble():
    from echoamd import Ensemble
    ic, rc, v0 = _imex(CASES["mhd-fireball-bjorken"]())
    a = _grid(ic, rc)
    b = _grid(ic, rc)
    c = _grid(ic, rc)
    for g in (a, b, c):
        assert g.init_from_prim(v0) == 0
    a.steps(3)
    for _ in range(3):
        for irk in range(1, 5):
            b.rk_stage(irk)
        b.end_step()
    c.set_timing(True)
    c.steps(3)
    va, ua = a.state()
    for g in (b, c):
        vb, ub = g.state()
        assert np.array_equal(va, vb) and np.array_equal(ua, ub)
    ic2, rc2, v2 = mhd2d(24, divclean=1)
    ens = Ensemble([(ic, rc), (ic2, rc2)])
    assert ens.grids[0].init_from_prim(v0) == 0 and ens.grids[1].init_from_prim(v2) == 0
    assert ens.steps(3) == 0
    ve, ue = ens.grids[0].state()
    assert np.array_equal(va, ve) and np.array_equal(ua, ue)
    ens.close()
    for g in (a, b, c):
        g.close()


def test_imex_refusals():
    from echoamd import C2pNet, EchoError
    from echoamd.surrogate import load_weights
    from scenarios import fireball3d
    ic, rc, v0 = _imex(fireball3d(12, 10, 8))
    with pytest.raises(EchoError, match="MHD grids"):
        _grid(ic, rc)
    ic, rc, v0 = _imex(mhd_fireball3d())
    g = _grid(ic, rc)
    net = C2pNet(0, load_weights())
    with pytest.raises(EchoError, match="IMEX"):
        g.set_c2p_net(net)
    net.close()
    g.close()
