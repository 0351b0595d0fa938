"""GPU parity: the HIP path (through the C-ABI) against the C oracle on the
same seeded inputs.  Bar: relative L-infinity <= 1e-10 in FP64 after N steps
(the tolerance the grid-update spec states for floating point)."""
import numpy as np
import pytest

from libs import Oracle
from scenarios import fireball3d, gubser2d, mhd2d, mhd3d, mhd_fireball3d, rel_linf, shock2d

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _gpu(ic, rc):
    from echoamd import EchoGrid
    return EchoGrid(ic, rc)


def _run_both(oracle, ic, rc, v0, nsteps, check_scalars=True):
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    g = _gpu(ic, rc)
    assert g.init_from_prim(v0) == 0
    vg0, ug0 = g.state()
    vo0, uo0 = oracle.state()
    assert rel_linf(vg0, vo0) <= TOL and rel_linf(ug0, uo0) <= TOL
    assert oracle.steps(nsteps) == 0
    g.steps(nsteps)
    vg, ug = g.state()
    vo, uo = oracle.state()
    so, sg = oracle.scalars(), g.scalars()
    if check_scalars:
        assert sg["iter"] == so["iter"] == nsteps
        assert abs(sg["t"] - so["t"]) <= TOL * abs(so["t"])
        assert abs(sg["dt"] - so["dt"]) <= TOL * abs(so["dt"])
    g.close()
    return (vg, ug), (vo, uo)


@pytest.mark.parametrize("rec", ["WENO5", "PPM4", "TVD2", "MPE5", "WENO3", "CENO3", "MPE3"])
def test_gubser_recon(oracle, rec):
    ic, rc, v0 = gubser2d(n=32, rec=rec)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 10)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


def test_gubser_cfg1_full(oracle):
    """Config 1: 128x128x1, WENO5 + HLL + RK3, CFL 0.4, 100 steps."""
    ic, rc, v0 = gubser2d(n=128)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 100)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


@pytest.mark.parametrize("eqstate", [1, 3])
@pytest.mark.parametrize("nrk,coords,bc", [(3, 2, 2), (2, 1, 0)])
def test_fireball_ideal(oracle, eqstate, nrk, coords, bc):
    ic, rc, v0 = fireball3d(eqstate=eqstate, nrk=nrk, coordinates=coords, bc=bc)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 8)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


@pytest.mark.parametrize("eqstate,smooth", [(1, 0.0), (3, 0.0), (3, 0.15)])
@pytest.mark.parametrize("coords", [2, 1])
def test_fireball_viscous(oracle, eqstate, smooth, coords):
    ic, rc, v0 = fireball3d(viscosity=1, eqstate=eqstate, smooth_temp=smooth, coordinates=coords)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 8)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


@pytest.mark.parametrize("rec", ["TVD2", "MPE5", "WENO3", "CENO3", "MPE3"])
def test_fireball_viscous_recon(oracle, rec):
    """The viscous flux sweep keeps the reconstructed pi components apart
    from rho, v, p: every reconstruction goes through that path."""
    ic, rc, v0 = fireball3d(viscosity=1, rec=rec)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 4)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


def test_fireball_viscous_no_zz(oracle):
    ic, rc, v0 = fireball3d(viscosity=1, obtained_zz=0, rec="PPM4", nrk=2)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 6)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


@pytest.mark.parametrize("rec", ["WENO5", "MPE5", "TVD2"])
def test_shock2d(oracle, rec):
    ic, rc, v0 = shock2d(rec=rec)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 20)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


def test_stagewise_api_matches_fused(oracle):
    """echo_rk_stage x nrk + echo_end_step == echo_steps (graph replay)."""
    from echoamd import EchoGrid
    ic, rc, v0 = gubser2d(n=32)
    a, b = EchoGrid(ic, rc), EchoGrid(ic, rc)
    a.init_from_prim(v0)
    b.init_from_prim(v0)
    for _ in range(5):
        for irk in (1, 2, 3):
            a.rk_stage(irk)
        a.end_step()
    b.steps(5)
    va, ua = a.state()
    vb, ub = b.state()
    assert np.array_equal(va, vb) and np.array_equal(ua, ub)


def test_graph_replay_interleaved_with_host_work(oracle):
    """The captured step graph replayed several times with host reads between
    replays gives the same state as the oracle after each replay."""
    ic, rc, v0 = fireball3d(nx=16, ny=12, nz=10)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    g = _gpu(ic, rc)
    g.init_from_prim(v0)
    for _ in range(4):
        g.steps(1)
        oracle.steps(1)
        vg, _ = g.state()
        vo, _ = oracle.state()
        assert rel_linf(vg, vo) <= TOL


def test_cons2prim_batch(oracle):
    rng = np.random.default_rng(7)
    ic, rc, v0 = fireball3d(eqstate=3)
    oracle.setup(ic, rc)
    oracle.set_time(1.0)
    g = _gpu(ic, rc)
    n = 4096
    v = np.zeros((5, n))
    v[0] = 1.0
    v[1:4] = rng.uniform(-0.5, 0.5, (3, n)) / np.sqrt(3)
    v[4] = np.exp(rng.uniform(np.log(0.05), np.log(5.0), n))
    u = np.zeros((5, n))
    for i in range(n):
        e, ui, _ = oracle.prim2cons(v[:, i])
        assert e == 0
        u[:, i] = ui
    guess = v.copy()
    guess[4] *= rng.uniform(0.8, 1.2, n)
    vg, st = g.cons2prim_batch(u, guess)
    assert (st == 0).all()
    for i in range(0, n, 97):
        e, vo, _ = oracle.cons2prim(u[:, i], guess[:, i])
        assert e == 0
        assert np.max(np.abs(vg[:, i] - vo) / np.maximum(np.abs(vo), 1e-300)) <= 1e-12 or \
            np.max(np.abs(vg[:, i] - vo)) <= 1e-14
    assert np.max(np.abs(vg - v) / np.maximum(np.abs(v), 1e-12)) < 1e-8


@pytest.mark.parametrize("algo,eqstate,coords,bc", [(5, 1, 1, 0), (5, 1, 2, 2), (6, 1, 1, 0), (4, 2, 1, 0)])
@pytest.mark.parametrize("divclean", [1, 0])
def test_mhd2d(oracle, algo, eqstate, coords, bc, divclean):
    ic, rc, v0 = mhd2d(32, eqstate=eqstate, algo=algo, divclean=divclean, coordinates=coords, bc=bc)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 12)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL


@pytest.mark.parametrize("divclean", [1, 0])
def test_mhd_fireball3d(oracle, divclean):
    ic, rc, v0 = mhd_fireball3d(divclean=divclean)
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, 6)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL
    assert abs(oracle.scalars()["glm_ch_old"] - _gpu_scalar_after(ic, rc, v0, 6)) <= 1e-10


def _gpu_scalar_after(ic, rc, v0, n):
    g = _gpu(ic, rc)
    g.init_from_prim(v0)
    g.steps(n)
    s = g.scalars()["glm_ch_old"]
    g.close()
    return s


@pytest.mark.parametrize("make,steps", [
    (lambda: mhd2d(32, eqstate=3, algo=1), 10),
    (lambda: mhd2d(32, eqstate=3, algo=2, divclean=0), 10),
    (lambda: mhd2d(32, algo=1, coordinates=2, bc=2), 10),
    (lambda: mhd3d(16, 14, 10, algo=1), 6),
    (lambda: mhd3d(16, 14, 10, algo=2), 6),
], ids=["2d-algo1-table", "2d-algo2-table-noglm", "2d-algo1-conformal-bjorken",
        "3d-algo1-table", "3d-algo2-table"])
def test_mhd_brentm(oracle, make, steps):
    """MHD con2prim algorithms 1/2 (BRENTM on the 3x3 / 2x2 systems) with the
    tabulated and conformal EoS."""
    ic, rc, v0 = make()
    (vg, ug), (vo, uo) = _run_both(oracle, ic, rc, v0, steps)
    assert rel_linf(vg, vo) <= TOL
    assert rel_linf(ug, uo) <= TOL
