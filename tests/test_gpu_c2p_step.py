"""The learned con2prim inside the grid's own steps (echo_set_c2p_net), ideal
MHD with p = e/3, through the C-ABI.

- A network whose guess the gate rejects everywhere hands every cell of
  every stage to the plain solver from the stage-start primitives, which is
  what the step does without a network: the states are bit-identical and
  the fallback count is cells x stages x steps.
- The trained network (weights/c2p_e3p_mlp.npy, 3 Newton iterations, gate
  1e-12) against the oracle's plain algorithm 5 after 5 steps, at the bar of
  the other trajectory tests (relative L-infinity 1e-10): the two inversions
  stop at different points inside their tolerances (rtsafe's 1e-12 in v^2,
  the gate's 1e-12 in W), far below the bar.
- Host-forced superluminal cells: the gate rejects them, the plain solver
  fails there, and the repair gives the oracle's count; the states within
  the oracle's own sensitivity to perturbations below rounding (this case
  amplifies rounding to ~1e-5 in two steps, in the oracle as on the GPU).
- In an ensemble the grid with a network equals its single grid bit for bit.
"""
import os

import numpy as np
import pytest

from scenarios import FORCED_CELLS, fireball3d, force_superluminal, mhd2d, mhd_fireball3d, rel_linf

pytestmark = pytest.mark.gpu
TOL = 1e-10

SCEN = {
    "bjorken-glm": lambda: mhd_fireball3d(),
    "minkowski": lambda: mhd2d(divclean=0),
    "minkowski-glm": lambda: mhd2d(divclean=1),
    "minkowski-glm-rk2": lambda: mhd2d(divclean=1, nrk=2),
}


def _grid(ic, rc):
    from echoamd import EchoGrid
    return EchoGrid(ic, rc)


def _trained():
    from echoamd import C2pNet
    from echoamd.surrogate import WEIGHTS, load_weights
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights (python -m echoamd.surrogate)")
    return C2pNet(0, load_weights())


def _rejecting():
    """A network whose W guess overflows (bias of the W head 1e4): the gate
    rejects every state."""
    from echoamd import C2pNet
    from echoamd.surrogate import load_weights
    P = load_weights().copy()
    P[9216 + 1] = 1e4
    return C2pNet(0, P)


@pytest.mark.parametrize("name", sorted(SCEN))
def test_rejecting_network_is_the_plain_step(name):
    ic, rc, v0 = SCEN[name]()
    print(f"\n{name}: {ic[:3]}", flush=True)
    a = _grid(ic, rc)
    b = _grid(ic, rc)
    assert a.init_from_prim(v0) == 0 and b.init_from_prim(v0) == 0
    net = _rejecting()
    b.set_c2p_net(net, iters=0, tol=1e-300)
    a.steps(5)
    b.steps(5)
    va, ua = a.state()
    vb, ub = b.state()
    assert np.array_equal(va, vb) and np.array_equal(ua, ub)
    n = va[0].size
    assert b.c2p_fallbacks() == n * int(ic[10]) * 5
    assert a.c2p_fallbacks() == 0
    a.close()
    b.close()
    net.close()


@pytest.mark.parametrize("name", sorted(SCEN))
def test_trained_network_in_step_vs_oracle(oracle, name):
    ic, rc, v0 = SCEN[name]()
    print(f"\n{name}: {ic[:3]}", flush=True)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(5) == 0
    vo, uo = oracle.state()
    net = _trained()
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.set_c2p_net(net, iters=3, tol=1e-12)
    g.steps(5)
    vg, ug = g.state()
    total = vg[0].size * int(ic[10]) * 5
    nfb = g.c2p_fallbacks()
    print(f"fallbacks {nfb} of {total}", flush=True)
    assert nfb < total
    assert g.scalars()["iter"] == 5
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL, (rel_linf(vg, vo), rel_linf(ug, uo))
    g.close()
    net.close()


def _oracle_forced(oracle, ic, rc, v0, pert=0.0, seed=0, steps=2):
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    v, u = oracle.state()
    u = force_superluminal(u)
    if pert:
        u = u * (1 + pert * np.random.default_rng(seed).standard_normal(u.shape))
    oracle.set_state(v, u)
    assert oracle.steps(steps) == 0
    vo, uo = oracle.state()
    return v, u, vo, uo, oracle.repair_counts()["u2v"]


def test_forced_failures_plain_and_learned(oracle):
    """Superluminal conserved states (S_x = 3 E) at three neighbouring cells
    of the ideal-MHD fireball, two steps, without and with the network: the
    repair runs on the cells the plain solver could not invert (with the
    network: after the gate rejected them), with the oracle's count.
    The states are compared against the oracle's own sensitivity: this case
    is ill-conditioned (the near-superluminal repaired cells and their
    reconstruction repairs amplify rounding), so the oracle run from inputs
    perturbed by 1e-16 relative, below one rounding, moves by ~1e-5 after two
    steps (against ~2e-14 without the forced cells).  Bar: the largest such
    spread over three perturbations, times 10."""
    ic, rc, v0 = mhd_fireball3d()
    v, u, vo, uo, n = _oracle_forced(oracle, ic, rc, v0)
    assert n >= len(FORCED_CELLS)
    spread = 0.0
    for seed in (1, 2, 3):
        _, _, vq, uq, nq = _oracle_forced(oracle, ic, rc, v0, pert=1e-16, seed=seed)
        assert nq == n
        spread = max(spread, rel_linf(vq, vo), rel_linf(uq, uo))
    print(f"\noracle spread under 1e-16 perturbations: {spread:.3e}", flush=True)
    net = _trained()
    out = {}
    for name in ("plain", "learned"):
        g = _grid(ic, rc)
        assert g.init_from_prim(v0) == 0
        g.upload(v, u)
        if name == "learned":
            g.set_c2p_net(net, iters=3, tol=1e-12)
        g.steps(2)
        vg, ug = g.state()
        out[name] = (g.repairs(), g.c2p_fallbacks())
        ev, eu = rel_linf(vg, vo), rel_linf(ug, uo)
        print(f"{name}: repairs {out[name][0]} (oracle {n}), fallbacks {out[name][1]}, v {ev:.3e} u {eu:.3e}",
              flush=True)
        g.close()
        assert out[name][0] == n, name
        assert ev <= 10 * spread and eu <= 10 * spread, (name, ev, eu, spread)
    net.close()
    assert out["plain"][1] == 0 and out["learned"][1] >= n


def test_network_grid_in_ensemble():
    from echoamd import Ensemble
    events = [mhd_fireball3d(), fireball3d(20, 16, 12, viscosity=1, eqstate=3)]
    net = _trained()
    ens = Ensemble([(ic, rc) for ic, rc, _ in events])
    for g, (ic, rc, v0) in zip(ens.grids, events):
        assert g.init_from_prim(v0) == 0
    ens.grids[0].set_c2p_net(net, iters=3, tol=1e-12)
    assert ens.steps(3) == 0
    ic, rc, v0 = events[0]
    s = _grid(ic, rc)
    assert s.init_from_prim(v0) == 0
    s.set_c2p_net(net, iters=3, tol=1e-12)
    s.steps(3)
    ve, ue = ens.grids[0].state()
    vs, us = s.state()
    assert np.array_equal(ve, vs) and np.array_equal(ue, us)
    assert ens.grids[0].c2p_fallbacks() == s.c2p_fallbacks()
    s.close()
    ens.close()
    net.close()


def test_detach_and_rejections():
    from echoamd import EchoError
    ic, rc, v0 = mhd_fireball3d()
    net = _trained()
    a = _grid(ic, rc)
    b = _grid(ic, rc)
    assert a.init_from_prim(v0) == 0 and b.init_from_prim(v0) == 0
    b.set_c2p_net(net)
    b.set_c2p_net(None)
    a.steps(3)
    b.steps(3)
    va, ua = a.state()
    vb, ub = b.state()
    assert np.array_equal(va, vb) and np.array_equal(ua, ub)
    with pytest.raises(EchoError, match="iters"):
        b.set_c2p_net(net, iters=9)
    a.close()
    b.close()
    for make in (lambda: fireball3d(12, 10, 8, viscosity=1, eqstate=3),
                 lambda: mhd_fireball3d(algo=1, eqstate=3)):
        ic, rc, v0 = make()
        g = _grid(ic, rc)
        with pytest.raises(EchoError, match="p = e/3"):
            g.set_c2p_net(net)
        g.close()
    net.close()
