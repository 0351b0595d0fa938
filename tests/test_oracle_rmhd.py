"""The IMEX-SSP3(4,3,3) integrator and the resistive MHD system (with and
without viscosity) in the oracle.

- IMEX-SSP3(4,3,3) (evolve.f90:127-130, 204-264): with ideal MHD the
  reference runs it (INT_ALGO = SSP is kept for MHD, common.f90:2076-2083,
  and the stiff part is zero without resistivity), so the oracle's stages
  are pinned bit for bit against the reference build.
- The resistive system has no working reference path (system.f90:1160-1162,
  common.f90:1904-1908); the oracle's restatement of the design
  (oracle/oracle_rmhd.c) is pinned by limits and closed forms:
  * without fields it is the hydro or viscous system bit for bit;
  * a uniform fluid at rest with a uniform E: E(t) = E0 exp(-sigma t) and
    the lost field energy reappears in the fluid (total energy exact);
  * an electromagnetic plane wave travels at the speed of light;
  * a uniform transverse field in Bjorken flow keeps tau B constant while
    the fluid follows the field-free Bjorken solution;
  * as sigma grows, E approaches the ideal Ohm's law E = -v x B (~1/sigma).
"""
import numpy as np
import pytest

from scenarios import (fireball3d, light_wave, mhd2d, mhd3d, mhd_fireball3d, rel_linf,
                       resistive_layout, rmhd_fireball, uniform_resistive)

HBARC = 0.197326

SSP_CASES = {
    "mhd2d-glm": lambda: mhd2d(24, divclean=1),
    "mhd2d": lambda: mhd2d(24, divclean=0),
    "mhd3d-brentm": lambda: mhd3d(10, 8, 6, algo=1),
    "mhd-fireball-bjorken": lambda: mhd_fireball3d(),
}


@pytest.mark.parametrize("name", sorted(SSP_CASES))
def test_imex_ssp_ideal_mhd_bitexact(oracle, ref, name):
    ic, rc, v0 = SSP_CASES[name]()
    ic = ic.copy()
    ic[10] = 4
    out = []
    for lib in (oracle, ref):
        lib.setup(ic, rc)
        assert lib.init_from_prim(v0) == 0
        assert lib.steps(4) == 0
        out.append(lib.state() + (lib.scalars(),))
    (vo, uo, so), (vr, ur, sr) = out
    assert so["t"] == sr["t"] and so["dt"] == sr["dt"] and so["glm_ch_old"] == sr["glm_ch_old"]
    assert np.array_equal(vo, vr) and np.array_equal(uo, ur)


def _run(oracle, ic, rc, v0, steps):
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(steps) == 0
    return oracle.state()


@pytest.mark.parametrize("viscosity", [0, 1])
def test_resistive_without_fields_is_the_fluid_system(oracle, viscosity):
    """E = B = psi = 0: the resistive system (RK3) equals the hydro or
    viscous system bit for bit (dt fixed by MAXDT in both, since their CFL
    speeds differ)."""
    ic, rc, v = rmhd_fireball(viscosity=viscosity, nrk=3, B0=0.0, E0=0.0, maxdt=0.02)
    kb = resistive_layout(viscosity)[0]
    vr, ur = _run(oracle, ic, rc, v, 4)
    icf, rcf, vf = fireball3d(16, 14, 10, viscosity=viscosity, nrk=3)
    rcf = rcf.copy()
    rcf[8] = 0.02
    vh, uh = _run(oracle, icf, rcf, vf, 4)
    assert np.array_equal(vr[:kb], vh) and np.array_equal(ur[:kb], uh)
    assert not vr[kb:].any() and not ur[kb:].any()


@pytest.mark.parametrize("nrk", [3, 4])
def test_uniform_field_decays_ohmically(oracle, nrk):
    """Uniform fluid at rest, uniform E, B = 0, Minkowski, sigma = 5.8 MeV:
    E(t) = E0 exp(-sigma t), sigma = 0.0058 / hbar c fm^-1, within the
    integrator's accuracy (sigma dt = 3e-4); the total energy of the cell is
    conserved exactly, so the fluid gains what the field loses; v stays 0."""
    E0 = np.array([0.3, -0.2, 0.1])
    ic, rc, v0 = uniform_resistive(E0=tuple(E0), nrk=nrk, maxdt=0.01)
    kb, ke, kg, nv = resistive_layout(0)
    nsteps = 50
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    u_start = oracle.state()[1]
    assert oracle.steps(nsteps) == 0
    v, u = oracle.state()
    t = oracle.scalars()["t"] - rc[6]
    sigma = 0.0058 / HBARC
    expect = E0 * np.exp(-sigma * t)
    for k in range(3):
        assert np.allclose(v[ke + k], expect[k], rtol=1e-10, atol=0)
    assert np.abs(u[4] - u_start[4]).max() <= 1e-14 * u_start[4].max()
    assert not v[1:4].any()
    # fluid energy gained = field energy lost (p = e/3)
    e_fluid = 3 * v[4]
    lost = 0.5 * (np.sum(E0 ** 2) - np.sum(expect ** 2))
    assert np.allclose(e_fluid - 3.0, lost, rtol=1e-8)


def test_stiff_relaxation_is_stable(oracle):
    """sigma dt = 100 (a near-ideal conductor): the IMEX stages damp E
    monotonically to zero instead of blowing up."""
    ic, rc, v0 = uniform_resistive(E0=(0.3, -0.2, 0.1), sigma_el=100 * HBARC / 0.01, nrk=4, maxdt=0.01)
    kb, ke, kg, nv = resistive_layout(0)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    prev = 0.3
    for _ in range(5):
        assert oracle.steps(1) == 0
        ex = abs(oracle.state()[0][ke].max())
        assert ex < prev
        prev = ex
    assert prev < 1e-3


@pytest.mark.parametrize("nrk", [3, 4])
def test_light_wave_travels_at_c(oracle, nrk):
    """E^y = B^eta = A sin(2 pi x) in a dense fluid, sigma = 0: after t = 0.25
    the profile is the initial one shifted by 0.25 (a wrong sign or factor in
    the Faraday or Ampere fluxes moves it elsewhere or not at all)."""
    ic, rc, v0 = light_wave(n=64, nrk=nrk)
    kb, ke, kg, nv = resistive_layout(0)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    oracle.c_set_tout = None
    t_end = 0.25
    while oracle.scalars()["t"] < t_end - 1e-12:
        assert oracle.steps(1) == 0
    t = oracle.scalars()["t"]
    v, _ = oracle.state()
    x = (np.arange(64) + 0.5) / 64 - 0.5
    exact = 1e-3 * np.sin(2 * np.pi * (x - t))
    assert np.abs(v[ke + 1, 0, 0] - exact).max() < 2e-3 * 1e-3
    assert np.abs(v[kb + 2, 0, 0] - exact).max() < 2e-3 * 1e-3


def test_bjorken_transverse_field(oracle):
    """Uniform fluid and uniform B^x in Bjorken coordinates (no E): tau B^x is
    conserved exactly, and the fluid follows the field-free run up to the
    scheme's first-order error in dt (the metric is advanced only after the
    whole step, as in the reference): halving dt halves the difference.  A
    wrong electromagnetic part of the Bjorken energy source would leave an
    O(1) difference that does not shrink."""
    kb, ke, kg, nv = resistive_layout(0)

    def fluid_p(B, dt, steps):
        ic, rc, v0 = uniform_resistive(E0=(0.0, 0.0, 0.0), B0=(B, 0.0, 0.0), p=2.0, coordinates=2, nrk=3,
                                       maxdt=dt)
        oracle.setup(ic, rc)
        assert oracle.init_from_prim(v0) == 0
        u0 = oracle.state()[1][kb].copy()
        assert oracle.steps(steps) == 0
        v, u = oracle.state()
        if B:
            assert np.abs(u[kb] - u0).max() <= 1e-15
        return v[4].mean(), oracle.scalars()["t"]

    diffs = []
    for dt, steps in ((0.02, 20), (0.01, 40)):
        pb, _ = fluid_p(0.5, dt, steps)
        p0, tau = fluid_p(0.0, dt, steps)
        diffs.append(abs(pb - p0) / p0)
    assert diffs[0] < 1e-3 and 0.4 < diffs[1] / diffs[0] < 0.6, diffs


def test_large_conductivity_approaches_ideal_ohm(oracle):
    """Growing sigma drives E towards -v x B: max |E + v x B| / |B| falls
    roughly as 1/sigma (ideal-MHD limit)."""
    res = []
    for sig in (2.0, 20.0, 200.0):
        ic, rc, v0 = rmhd_fireball(viscosity=0, nrk=4, B0=0.3, E0=0.0, coordinates=1, sigma_el=sig * HBARC,
                                   maxdt=0.01)
        kb, ke, kg, nv = resistive_layout(0)
        oracle.setup(ic, rc)
        assert oracle.init_from_prim(v0) == 0
        assert oracle.steps(30) == 0
        v, _ = oracle.state()
        vel, B, E = v[1:4], v[kb:kb + 3], v[ke:ke + 3]
        vxb = np.cross(vel, B, axis=0)
        res.append(np.abs(E + vxb).max() / np.abs(B).max())
    assert res[1] < 0.3 * res[0] and res[2] < 0.3 * res[1], res
