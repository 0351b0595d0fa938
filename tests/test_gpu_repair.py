"""GPU parity of the repair branches: the reconstruction repair inside k_flux
(evolve.f90:334-369) and the con2prim repair of k_update + k_fixup
(evolve_u2v, evolve.f90:795-842), through the C-ABI, against the oracle on
the states of tests/test_oracle_repair.py (where the oracle equals the
reference bit for bit).  Bar: relative L-infinity <= 1e-10 after the steps,
and the same number of repaired cells.

The GPU repairs the failed cells of a stage in dependency levels (a failed
cell waits for its failed face neighbours that come earlier in the
reference's loop), reading the neighbours as the reference's sequential loop
leaves them; the forced cases make neighbouring cells fail together, so a
wrong order or a wrong reading shows up at order one.
"""
import numpy as np
import pytest

from scenarios import FORCED_CELLS, fast_eta, force_superluminal, forced_background, rel_linf, rough_patch

pytestmark = pytest.mark.gpu
TOL = 1e-10

ROUGH = {
    "rough-ideal-weno5": lambda: rough_patch(eqstate=1, rec="WENO5", seed=1),
    "rough-ideal-mpe5": lambda: rough_patch(eqstate=1, rec="MPE5", seed=0),
    "rough-table-weno5": lambda: rough_patch(eqstate=3, rec="WENO5", seed=0, pr_lo=-3.0),
}
FORCED = {
    "forced-idealgas-tvd2": dict(eqstate=2, rec="TVD2", pbg=1e-3),
    "forced-idealgas-weno5": dict(eqstate=2, rec="WENO5", pbg=1e-2),
    "forced-viscous-weno5": dict(eqstate=1, viscosity=1, rec="WENO5", pbg=1.0),
    "forced-viscous-smooth": dict(eqstate=1, viscosity=1, rec="WENO5", pbg=0.3, smooth_temp=0.15),
}


def _grid(ic, rc):
    from echoamd import EchoGrid
    return EchoGrid(ic, rc)


def _forced(oracle, kw):
    ic, rc, v0 = forced_background(**kw)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    v, u = oracle.state()
    return ic, rc, v0, v, force_superluminal(u)


@pytest.mark.parametrize("name", sorted(ROUGH))
def test_rough_patch_repairs(oracle, name):
    ic, rc, v0 = ROUGH[name]()
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(3) == 0
    cnt = oracle.repair_counts()
    assert cnt["rec_pr"] > 0 and cnt["rec_v"] > 0 and cnt["u2v"] > 0
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.steps(3)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert g.repairs() == cnt["u2v"]
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL
    g.close()


@pytest.mark.parametrize("name", sorted(FORCED))
def test_forced_failures(oracle, name):
    ic, rc, v0, v, u = _forced(oracle, FORCED[name])
    oracle.set_state(v, u)
    assert oracle.steps(2) == 0
    cnt = oracle.repair_counts()
    assert cnt["u2v"] >= len(FORCED_CELLS)
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.upload(v, u)
    g.steps(2)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert g.repairs() == cnt["u2v"]
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL
    g.close()


def test_failed_repair_crash(oracle):
    """fast_eta: step 1 repairs cells, step 2 rescales two cells to maxspeed
    in the reconstruction repair and then a con2prim repair fails.  The GPU
    reports the reference's error code (1) with a message naming the crashing
    cell, which is the oracle's; the cells before it in the reference's loop
    match the oracle (the reference stops its loop at the crash)."""
    from echoamd import EchoError
    ic, rc, v0 = fast_eta(rec="TVD2")
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(1) == 0
    n1 = oracle.repair_counts()["u2v"]
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.steps(1)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert g.repairs() == n1
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL
    assert oracle.steps(1) == 1
    cnt = oracle.repair_counts()
    assert cnt["rec_vmax"] > 0 and cnt["u2v_fail"] == 1
    crash = oracle.repair_cells()[cnt["u2v"] - 1]
    with pytest.raises(EchoError) as ei:
        g.steps(1)
    assert ei.value.code == 1
    nx, ny = int(ic[0]), int(ic[1])
    ix, iy, iz = crash % nx + 1, (crash // nx) % ny + 1, crash // (nx * ny) + 1
    assert f"ix={ix} iy={iy} iz={iz}" in str(ei.value), str(ei.value)
    s = g.scalars()
    assert s["crashed"] == 1 and s["errcode"] == 1
    vg, ug = g.state()
    vo, uo = oracle.state()
    nv = vg.shape[0]
    before = np.arange(vg[0].size) < crash
    a = vg.reshape(nv, -1)[:, before].reshape(nv, 1, 1, -1)
    b = vo.reshape(nv, -1)[:, before].reshape(nv, 1, 1, -1)
    assert rel_linf(a, b) <= TOL
    g.close()


def test_repairs_in_ensemble(oracle):
    """Two events with failures in one ensemble (one shared set of repair
    scratch buffers): each matches its own grid bit for bit and the oracle."""
    from echoamd import Ensemble
    e1 = _forced(oracle, FORCED["forced-idealgas-tvd2"])
    ic2, rc2, v02 = ROUGH["rough-ideal-weno5"]()
    ens = Ensemble([(e1[0], e1[1]), (ic2, rc2)])
    g1, g2 = ens.grids
    assert g1.init_from_prim(e1[2]) == 0
    g1.upload(e1[3], e1[4])
    assert g2.init_from_prim(v02) == 0
    assert ens.steps(2) == 0
    oracle.setup(e1[0], e1[1])
    oracle.init_from_prim(e1[2])
    oracle.set_state(e1[3], e1[4])
    oracle.steps(2)
    vo, uo = oracle.state()
    v1, u1 = g1.state()
    assert g1.repairs() == oracle.repair_counts()["u2v"] > 0
    assert rel_linf(v1, vo) <= TOL and rel_linf(u1, uo) <= TOL
    oracle.setup(ic2, rc2)
    oracle.init_from_prim(v02)
    oracle.steps(2)
    vo, uo = oracle.state()
    v2, u2 = g2.state()
    assert g2.repairs() == oracle.repair_counts()["u2v"] > 0
    assert rel_linf(v2, vo) <= TOL and rel_linf(u2, uo) <= TOL
    ens.close()


# a 4 x 4 x 4 block of superluminal cells: chains of up to ten dependent
# repairs per stage (levels 0..9), where each cell waits for its failed
# neighbours at x-1, y-1, z-1 (the oracle's spread under 1e-16 perturbations
# of these inputs is 2e-13)
BLOCK = tuple((z, y, x) for z in range(2, 6) for y in range(3, 7) for x in range(3, 7))


@pytest.mark.parametrize("name", ["forced-idealgas-tvd2", "forced-viscous-weno5"])
def test_forced_block_chains(oracle, name):
    ic, rc, v0 = forced_background(**FORCED[name])
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    v, u = oracle.state()
    u = force_superluminal(u, cells=BLOCK)
    oracle.set_state(v, u)
    assert oracle.steps(2) == 0
    cnt = oracle.repair_counts()
    assert cnt["u2v"] >= len(BLOCK)
    print(f"\n{name}: {cnt['u2v']} repairs", flush=True)
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.upload(v, u)
    g.steps(2)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert g.repairs() == cnt["u2v"]
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL, (rel_linf(vg, vo), rel_linf(ug, uo))
    g.close()
