"""GPU parity of the resistive MHD systems (MHD = 2: nv = 12, and nv = 19
with viscosity) and of the IMEX-SSP3(4,3,3) integrator (INT_ALGO 4),
through the C-ABI, against the oracle on the same seeded inputs.

The oracle's IMEX stages are pinned bit for bit against the reference build
with ideal MHD, and its resistive terms by limits and closed forms
(tests/test_oracle_rmhd.py).  Bar: relative L-infinity <= 1e-10 in FP64
after the steps; for the forced failures also the same number of repaired
cells (the repairs inside the implicit stages read the earlier cells with
their solved E and the later ones with the stage start, as the oracle's
sequential loop does)."""
import numpy as np
import pytest

from scenarios import (force_superluminal, forced_resistive, light_wave, resistive_layout, rel_linf,
                       rmhd_fireball, uniform_resistive)

pytestmark = pytest.mark.gpu
TOL = 1e-10

FIREBALL = {
    "rmhd-rk3": dict(viscosity=0, nrk=3),
    "rmhd-imex": dict(viscosity=0, nrk=4),
    "vrmhd-rk3": dict(viscosity=1, nrk=3),
    "vrmhd-imex": dict(viscosity=1, nrk=4),
    "vrmhd-imex-minkowski": dict(viscosity=1, nrk=4, coordinates=1),
    "vrmhd-imex-nozz": dict(viscosity=1, nrk=4, obtained_zz=0),
    "vrmhd-imex-smooth": dict(viscosity=1, nrk=4, smooth_temp=0.15),
    "vrmhd-imex-stiff": dict(viscosity=1, nrk=4, sigma_el=50.0),
    "vrmhd-imex-mpe5": dict(viscosity=1, nrk=4, rec="MPE5"),
    "vrmhd-imex-ppm4": dict(viscosity=1, nrk=4, rec="PPM4"),
}


def _grid(ic, rc):
    from echoamd import EchoGrid
    return EchoGrid(ic, rc)


def _compare(oracle, g, steps):
    vg, ug = g.state()
    vo, uo = oracle.state()
    so, sg = oracle.scalars(), g.scalars()
    assert sg["iter"] == so["iter"] == steps
    assert abs(sg["t"] - so["t"]) <= TOL * abs(so["t"])
    assert abs(sg["dt"] - so["dt"]) <= TOL * abs(so["dt"])
    assert sg["glm_ch_old"] == so["glm_ch_old"] == 1.0
    return rel_linf(vg, vo), rel_linf(ug, uo)


@pytest.mark.parametrize("name", sorted(FIREBALL))
def test_resistive_fireball(oracle, name):
    ic, rc, v0 = rmhd_fireball(**FIREBALL[name])
    print(f"\n{name}: {ic[:3]} nv {v0.shape[0]}", flush=True)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    g = _grid(ic, rc)
    assert g.nv == v0.shape[0]
    assert g.init_from_prim(v0) == 0
    assert oracle.steps(5) == 0
    g.steps(5)
    ev, eu = _compare(oracle, g, 5)
    assert ev <= TOL and eu <= TOL, (ev, eu)
    # the fields moved: E relaxes, the current flows
    v, _ = g.state()
    kb, ke, kg, nv = resistive_layout(FIREBALL[name]["viscosity"])
    assert np.abs(v[ke:ke + 3] - v0[ke:ke + 3]).max() > 1e-4
    g.close()


@pytest.mark.parametrize("viscosity", [0, 1])
@pytest.mark.parametrize("nrk", [3, 4])
def test_resistive_forced_repairs(oracle, viscosity, nrk):
    """Three neighbouring cells with superluminal conserved states: every
    inversion of the first step fails there (five per IMEX step: K1..K4 and
    the result) and is repaired."""
    ic, rc, v0 = forced_resistive(viscosity=viscosity, nrk=nrk)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    v, u = oracle.state()
    u = force_superluminal(u)
    oracle.set_state(v, u)
    assert oracle.steps(2) == 0
    n = oracle.repair_counts()["u2v"]
    assert n == (15 if nrk == 4 else 9)
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.upload(v, u)
    g.steps(2)
    assert g.repairs() == n
    ev, eu = _compare(oracle, g, 2)
    assert ev <= TOL and eu <= TOL, (ev, eu)
    g.close()


@pytest.mark.parametrize("nrk", [3, 4])
def test_ohmic_decay_closed_form(nrk):
    """Uniform fluid at rest with a uniform E: E(t) = E0 exp(-sigma t) on the
    GPU as in the oracle (sigma = 5.8 MeV / hbar c)."""
    E0 = np.array([0.3, -0.2, 0.1])
    ic, rc, v0 = uniform_resistive(E0=tuple(E0), nrk=nrk, maxdt=0.01)
    kb, ke, kg, nv = resistive_layout(0)
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.steps(50)
    v, _ = g.state()
    t = g.scalars()["t"] - rc[6]
    expect = E0 * np.exp(-0.0058 / 0.197326 * t)
    for k in range(3):
        assert np.allclose(v[ke + k], expect[k], rtol=1e-10, atol=0)
    assert not v[1:4].any()
    g.close()


@pytest.mark.parametrize("nrk", [3, 4])
def test_light_wave(oracle, nrk):
    """The plane wave of test_oracle_rmhd.py.  The fluid stays at rest: its
    velocity (~2e-12) is the rounding left by subtracting the Poynting flux
    from the total momentum, so it is compared in absolute terms (1e-12 of
    the speed of light); everything else at the relative bar."""
    ic, rc, v0 = light_wave(n=64, nrk=nrk)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(20) == 0
    g = _grid(ic, rc)
    assert g.init_from_prim(v0) == 0
    g.steps(20)
    _compare(oracle, g, 20)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert np.abs(vg[1:4] - vo[1:4]).max() <= 1e-12
    vg[1:4] = vo[1:4] = 0.
    assert rel_linf(vg, vo) <= TOL
    # u: the momenta (~ the Poynting flux, 1e-6) against the energy scale
    for k in range(ug.shape[0]):
        d = np.abs(ug[k] - uo[k]).max()
        if k in (1, 2, 3):
            assert d <= 1e-12 * np.abs(uo[4]).max(), k
        else:
            assert d <= TOL * np.abs(uo[k]).max(), k
    g.close()


def test_stages_one_at_a_time():
    """echo_rk_stage 1..4 + echo_end_step give the captured step's result
    (the IMEX launches are grouped per stage the same way)."""
    ic, rc, v0 = rmhd_fireball(viscosity=1, nrk=4)
    a = _grid(ic, rc)
    b = _grid(ic, rc)
    assert a.init_from_prim(v0) == 0 and b.init_from_prim(v0) == 0
    a.steps(2)
    for _ in range(2):
        for irk in range(1, 5):
            b.rk_stage(irk)
        b.end_step()
    va, ua = a.state()
    vb, ub = b.state()
    assert np.array_equal(va, vb) and np.array_equal(ua, ub)
    a.close()
    b.close()


def test_unsupported_options_rejected():
    from echoamd import EchoError
    ic, rc, v0 = rmhd_fireball(viscosity=1, nrk=4)
    bad = ic.copy()
    bad[7] = 0   # no divergence cleaning
    with pytest.raises(EchoError, match="DIVCLEAN"):
        _grid(bad, rc)


def test_resistive_events_in_ensemble(oracle):
    """Resistive grids (nv 12 and 19, IMEX and RK3, their flux buffers with
    the extra div E slot) in one ensemble with a viscous-hydro grid, on the
    shared face pool: each equals its own single grid bit for bit, and the
    viscous resistive one the oracle."""
    from echoamd import EchoGrid, Ensemble
    from scenarios import fireball3d
    events = [rmhd_fireball(viscosity=1, nrk=4), rmhd_fireball(12, 10, 8, viscosity=0, nrk=3),
              fireball3d(20, 16, 12, viscosity=1, eqstate=3)]
    ens = Ensemble([(ic, rc) for ic, rc, _ in events])
    for g, (ic, rc, v0) in zip(ens.grids, events):
        assert g.init_from_prim(v0) == 0
    assert ens.steps(3) == 0
    for i, (g, (ic, rc, v0)) in enumerate(zip(ens.grids, events)):
        ve, ue = g.state()
        s = EchoGrid(ic, rc)
        s.init_from_prim(v0)
        s.steps(3)
        vs, us = s.state()
        s.close()
        assert np.array_equal(ve, vs) and np.array_equal(ue, us), i
    ic, rc, v0 = events[0]
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(3) == 0
    vo, uo = oracle.state()
    ve, ue = ens.grids[0].state()
    assert rel_linf(ve, vo) <= TOL and rel_linf(ue, uo) <= TOL
    ens.close()
