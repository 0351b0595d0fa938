"""GPU checks of the flux tile shapes, the autotuner and the event ensemble.

- every tile shape of the y and z sweeps gives bit-identical states (the
  per-face arithmetic does not depend on how faces are grouped in blocks);
- the autotuner leaves the state unchanged and picks a valid shape;
- an ensemble of uneven events (different sizes, physics and EoS) advanced
  by one captured graph per step equals each event run on its own grid, bit
  for bit, and the oracle to 1e-10.
"""
import numpy as np
import pytest

from scenarios import fireball3d, mhd2d, mhd_fireball3d, rel_linf

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _grid(ic, rc):
    from echoamd import EchoGrid
    return EchoGrid(ic, rc)


@pytest.mark.parametrize("make", [
    lambda: fireball3d(24, 20, 16, viscosity=1, eqstate=3),
    lambda: fireball3d(20, 18, 12, rec="PPM4"),
    lambda: mhd_fireball3d(),
], ids=["visc-tab-weno5", "ideal-ppm4", "mhd-glm-weno5"])
def test_tile_shapes_identical(make):
    ic, rc, v0 = make()
    ref = None
    for ty in (1, 2, 4):
        g = _grid(ic, rc)
        g.set_tile(1, ty)
        g.set_tile(2, ty)
        assert list(g.tile()) == [1, ty, ty]
        assert g.init_from_prim(v0) == 0
        g.steps(5)
        v, u = g.state()
        g.close()
        if ref is None:
            ref = (v, u)
        else:
            assert np.array_equal(v, ref[0]) and np.array_equal(u, ref[1]), ty


def test_tile_rejected_for_other_reconstructions():
    from echoamd import EchoError
    ic, rc, v0 = fireball3d(12, 10, 8, rec="MPE5")
    g = _grid(ic, rc)
    with pytest.raises(EchoError):
        g.set_tile(1, 2)
    with pytest.raises(EchoError):
        g.set_tile(0, 2)
    g.close()


def test_tune_keeps_state(oracle):
    ic, rc, v0 = fireball3d(32, 28, 20, viscosity=1, eqstate=3)
    g = _grid(ic, rc)
    g.init_from_prim(v0)
    g.steps(2)
    before = g.state()
    ms = g.tune(reps=3)
    after = g.state()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    t = g.tile()
    assert t[0] == 1 and t[1] in (1, 2, 4) and t[2] in (1, 2, 4)
    assert np.all(ms[1:] > 0)
    print(f"\ntuned tiles {t.tolist()}, ms per launch (rows 1/2/4): y {ms[1]} z {ms[2]}")
    # the tuned grid continues exactly like the oracle
    g.steps(3)
    vg, ug = g.state()
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(5)
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL
    g.close()


def _events():
    return [fireball3d(24, 20, 16, viscosity=1, eqstate=3),
            fireball3d(16, 12, 10),
            mhd2d(32),
            fireball3d(20, 16, 12, viscosity=1, eqstate=3, seed=3)]


@pytest.mark.parametrize("tune", [False, True])
def test_ensemble_matches_single_grids(oracle, tune):
    from echoamd import Ensemble
    events = _events()
    ens = Ensemble([(ic, rc) for ic, rc, _ in events])
    assert len(ens) == len(events)
    for g, (ic, rc, v0) in zip(ens.grids, events):
        assert g.init_from_prim(v0) == 0
    if tune:
        ens.tune(reps=2)
    nsteps = 4
    assert ens.steps(nsteps) == 0
    assert ens.last_ms() > 0
    for i, (g, (ic, rc, v0)) in enumerate(zip(ens.grids, events)):
        ve, ue = g.state()
        s = _grid(ic, rc)
        s.init_from_prim(v0)
        s.steps(nsteps)
        vs, us = s.state()
        s.close()
        assert np.array_equal(ve, vs) and np.array_equal(ue, us), i
        assert g.scalars()["iter"] == nsteps
    # one event against the oracle
    ic, rc, v0 = events[0]
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(nsteps)
    vo, uo = oracle.state()
    ve, ue = ens.grids[0].state()
    assert rel_linf(ve, vo) <= TOL and rel_linf(ue, uo) <= TOL
    ens.close()
