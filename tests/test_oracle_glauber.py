"""The oracle's Glauber-MC initial state (oracle/oracle_init.c) pinned by
closed forms: a single source is the normalised Gaussian of
glaubermc.f90:838 times the longitudinal profile; a cloud of narrow sources
integrates to kappa per (weighted) source; swapping the nuclei mirrors the
state in eta; the pressure is the EoS at e + enezero; the viscous stresses
are the Navier-Stokes values of viscous_initio."""
import numpy as np
import pytest

from libs import HBAR, cell_centres, make_config, oracle_glauber_prim

PAR = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=5.36, deta=3.0, sigeta=0.8, rhocenter=1.0)


def _setup(oracle, nx=40, ny=36, nz=9, viscosity=0, eqstate=1, zlim=(-4.5, 4.5), half=4.0):
    ic, rc = make_config(nx=nx, ny=ny, nz=nz, coordinates=2, viscosity=viscosity, eqstate=eqstate,
                         xlim=(-half, half), ylim=(-half * ny / nx, half * ny / nx), zlim=zlim,
                         tmin=0.6, init_type=3, eta_over_s=0.1)
    oracle.setup(ic, rc)
    return ic, rc


def _feta(z, p):
    h = p["deta"] / 2
    return np.where(np.abs(z) >= p["ybeam"], 0.0,
                    np.where(np.abs(z) <= h, 1.0, np.exp(-(np.abs(z) - h) ** 2 / (2 * p["sigeta"] ** 2))))


def test_single_source_closed_form(oracle):
    ic, rc = _setup(oracle)
    p = dict(PAR, ah=0.0, rhocenter=1.0)
    e, v = oracle_glauber_prim(oracle, p, [[0.7, -0.3]], np.zeros((0, 2)), np.zeros((0, 2)))
    assert e == 0
    nx, ny, nz = ic[0], ic[1], ic[2]
    x = cell_centres(nx, rc[0], rc[1])
    y = cell_centres(ny, rc[2], rc[3])
    z = cell_centres(nz, rc[4], rc[5])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    s2 = p["sig_mc"] ** 2
    # the source sits at the centre of mass, i.e. at the origin
    edens = p["kappa"] / (2 * np.pi * s2) * np.exp(-(X ** 2 + Y ** 2) / (2 * s2)) \
        * (p["ybeam"] + Z) / p["ybeam"] * _feta(Z, p)
    pr = (edens + rc[12]) / 3.0          # p = e/3 above przero
    assert np.allclose(v[4], pr, rtol=1e-13, atol=1e-300)
    assert np.all(v[0] == 1.0) and np.all(v[1:4] == 0.0)


def test_normalisation_and_mirror(oracle):
    rng = np.random.default_rng(3)
    ic, rc = _setup(oracle, nx=96, ny=96, nz=5, zlim=(-1.0, 1.0), half=6.0)
    pa = rng.normal(0, 1.2, (60, 2))
    pb = rng.normal(0, 1.2, (55, 2))
    pc = rng.normal(0, 1.0, (140, 2))
    e, v = oracle_glauber_prim(oracle, PAR, pa, pb, pc)
    assert e == 0
    dx = (rc[1] - rc[0]) / ic[0]
    dy = (rc[3] - rc[2]) / ic[1]
    # central eta slice (|eta| < deta/2, weights (ybeam +- eta)/ybeam sum to 2 per pair)
    iz = ic[2] // 2
    z0 = cell_centres(ic[2], rc[4], rc[5])[iz]
    edens = 3.0 * v[4, iz] - rc[12]
    tot = edens.sum() * dx * dy
    expect = PAR["kappa"] * ((1 - PAR["ah"]) * (60 * (PAR["ybeam"] + z0) / PAR["ybeam"]
                                                + 55 * (PAR["ybeam"] - z0) / PAR["ybeam"])
                             + PAR["ah"] * 140)
    assert abs(tot / expect - 1) < 1e-6
    # swapping the nuclei mirrors the state in eta
    e2, w = oracle_glauber_prim(oracle, PAR, pb, pa, pc)
    assert e2 == 0
    assert np.allclose(w[4], v[4, ::-1], rtol=1e-13)


def test_viscous_navier_stokes_stresses(oracle):
    rng = np.random.default_rng(4)
    ic, rc = _setup(oracle, nx=24, ny=20, nz=6, viscosity=1, eqstate=3)
    pa, pb, pc = rng.normal(0, 1.5, (30, 2)), rng.normal(0, 1.5, (30, 2)), rng.normal(0, 1.2, (50, 2))
    e, v = oracle_glauber_prim(oracle, PAR, pa, pb, pc)
    assert e == 0
    tau0 = rc[6]
    for idx in [(0, 3, 5), (3, 10, 12), (5, 19, 23)]:
        pr = v[(4,) + idx]
        err, out = oracle.eos_derived(1.0, pr)   # (en, temp, s)
        assert err == 0
        eta = HBAR * rc[10] * out[2]
        assert np.isclose(v[(10,) + idx], 2 * eta / (3 * tau0), rtol=1e-14)   # pi^yy
        assert np.isclose(v[(9,) + idx], 2 * eta / (3 * tau0), rtol=1e-14)    # pi^xx
        assert np.isclose(v[(11,) + idx], -tau0 ** 2 * 4 * eta / (3 * tau0), rtol=1e-14)
        assert v[(5,) + idx] == 0.0                                           # no bulk
