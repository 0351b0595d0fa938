"""The C oracle against the reference Fortran (oracle/_ref, built from the
reference sources): function level and trajectories, bit for bit.

Skipped where oracle/_ref/libechoref.so is not built (the reference sources
are needed to build it); tests/test_oracle_golden.py then checks the oracle
against fixtures written by the same library."""
import numpy as np
import pytest

from libs import REC_NGC, make_config
from scenarios import fireball3d, gubser2d, mhd2d, mhd3d, mhd_fireball3d, shock2d

RNG = np.random.default_rng(20261202)


@pytest.mark.parametrize("code", sorted(REC_NGC))
def test_reconstruction_bitexact(oracle, ref, code):
    ngc = REC_NGC[code]
    for n in (1, 5, 17):
        for kind in ("smooth", "noise", "step"):
            x = np.arange(n + 2 * ngc, dtype=float)
            if kind == "smooth":
                f = np.sin(0.3 * x) + 0.1 * x
            elif kind == "noise":
                f = RNG.standard_normal(x.size)
            else:
                f = np.where(x < x.size / 2, 1.0, 0.125)
            gl_o, gr_o = oracle.rec(code, f)
            gl_r, gr_r = ref.rec(code, f)
            assert np.array_equal(gl_o, gl_r) and np.array_equal(gr_o, gr_r), (code, n, kind)


@pytest.mark.parametrize("eqstate", [1, 2, 3])
def test_eos_bitexact(oracle, ref, eqstate):
    ic, rc = make_config(eqstate=eqstate)
    oracle.setup(ic, rc)
    ref.setup(ic, rc)
    assert oracle.c_przero() == ref.c_przero()
    if eqstate == 3:
        to, tr = oracle.eos_table(), ref.eos_table()
        for a, b in zip(to, tr):
            assert np.array_equal(a, b)
    en = np.exp(RNG.uniform(np.log(1e-4), np.log(3000.0), 400))
    pr = en / 3.0 * RNG.uniform(0.5, 1.0, en.size)
    for e, p in zip(en, pr):
        rh = 1.0 + RNG.uniform()
        assert oracle.eos_pressure(rh, e)[1].tolist() == ref.eos_pressure(rh, e)[1].tolist()
        assert oracle.eos_energy(rh, p) == ref.eos_energy(rh, p)
        assert oracle.eos_sound(rh, e, p) == ref.eos_sound(rh, e, p)
        assert oracle.eos_derived(rh, p)[1].tolist() == ref.eos_derived(rh, p)[1].tolist()


def _random_prim(nv, n):
    v = np.zeros((n, nv))
    v[:, 0] = RNG.uniform(0.5, 2.0, n)
    v[:, 1:4] = RNG.uniform(-0.5, 0.5, (n, 3)) / np.sqrt(3)
    v[:, 4] = np.exp(RNG.uniform(np.log(0.01), np.log(50.0), n))
    if nv == 12:
        v[:, 5:12] = RNG.uniform(-0.05, 0.05, (n, 7)) * v[:, 4:5]
    return v


@pytest.mark.parametrize("viscosity", [0, 1])
@pytest.mark.parametrize("eqstate", [1, 3])
@pytest.mark.parametrize("coords,t", [(1, 1.0), (2, 0.6), (2, 3.0)])
def test_pointwise_system_bitexact(oracle, ref, viscosity, eqstate, coords, t):
    ic, rc = make_config(viscosity=viscosity, eqstate=eqstate, coordinates=coords)
    oracle.setup(ic, rc)
    ref.setup(ic, rc)
    oracle.set_time(t)
    ref.set_time(t)
    nv = oracle.nv
    for v in _random_prim(nv, 60):
        e1, u1, v1 = oracle.prim2cons(v)
        e2, u2, v2 = ref.prim2cons(v)
        assert e1 == e2 and np.array_equal(u1, u2) and np.array_equal(v1, v2)
        for d in (1, 2, 3):
            a = oracle.flux(v, d)
            b = ref.flux(v, d)
            assert a[0] == b[0]
            for x, y in zip(a[1:], b[1:]):
                assert np.array_equal(x, y)
        guess = v.copy()
        guess[4] *= 1.1
        if viscosity:
            guess[1:4] *= 0.9
        a = oracle.cons2prim(u1, guess)
        b = ref.cons2prim(u1, guess)
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


SCENARIOS = {
    "gubser-weno5-rk3": (lambda: gubser2d(24), 12),
    "gubser-ppm4-rk2": (lambda: gubser2d(24, rec="PPM4", nrk=2), 12),
    "gubser-mpe5": (lambda: gubser2d(24, rec="MPE5"), 8),
    "shock-ideal-gas": (lambda: shock2d(32), 15),
    "fireball-ideal-tab": (lambda: fireball3d(12, 10, 8, eqstate=3), 5),
    "fireball-visc-tab-smooth": (lambda: fireball3d(12, 10, 8, viscosity=1, eqstate=3,
                                                    smooth_temp=0.15), 5),
    "fireball-visc-mink-periodic": (lambda: fireball3d(12, 10, 8, viscosity=1, coordinates=1,
                                                       bc=0, nrk=2), 5),
    "fireball-visc-no-zz": (lambda: fireball3d(12, 10, 8, viscosity=1, obtained_zz=0), 5),
    "fireball-visc-tab": (lambda: fireball3d(12, 10, 8, viscosity=1, eqstate=3), 5),
    "mhd2d-algo5-glm": (lambda: mhd2d(24), 8),
    "mhd2d-algo5-noglm-ppm4": (lambda: mhd2d(24, divclean=0, rec="PPM4"), 8),
    "mhd2d-algo6": (lambda: mhd2d(24, algo=6), 8),
    "mhd2d-algo4-idealgas": (lambda: mhd2d(24, eqstate=2, algo=4), 8),
    "mhd2d-bjorken-outflow": (lambda: mhd2d(24, coordinates=2, bc=2), 8),
    "mhd-fireball3d-glm": (lambda: mhd_fireball3d(12, 10, 8), 4),
    "mhd-fireball3d-noglm-rk2": (lambda: mhd_fireball3d(12, 10, 8, divclean=0), 4),
    "mhd2d-algo1-table": (lambda: mhd2d(24, eqstate=3, algo=1), 6),
    "mhd2d-algo2-table-noglm": (lambda: mhd2d(24, eqstate=3, algo=2, divclean=0), 6),
    "mhd2d-algo1-conformal-bjorken": (lambda: mhd2d(24, algo=1, coordinates=2, bc=2), 6),
    "mhd3d-algo1-table-bjorken": (lambda: mhd3d(12, 10, 8, algo=1), 5),
    "mhd3d-algo2-table-bjorken": (lambda: mhd3d(12, 10, 8, algo=2), 5),
}


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_trajectory_bitexact(oracle, ref, name):
    make, n = SCENARIOS[name]
    ic, rc, v0 = make()
    for lib in (oracle, ref):
        lib.setup(ic, rc)
        assert lib.init_from_prim(v0) == 0
    vo, uo = oracle.state()
    vr, ur = ref.state()
    assert np.array_equal(vo, vr) and np.array_equal(uo, ur)
    assert oracle.steps(n) == 0 and ref.steps(n) == 0
    vo, uo = oracle.state()
    vr, ur = ref.state()
    so, sr = oracle.scalars(), ref.scalars()
    for k in ("t", "dt", "timeold", "timeinterval", "g_cov_3_old", "tout", "iter", "glm_ch_old"):
        assert so[k] == sr[k], k
    assert np.array_equal(vo, vr)
    # the viscous smoothing threshold comes from a bisection that the
    # reference starts from an uninitialised bound: agreement to 1e-13 there
    if "smooth" in name:
        assert np.max(np.abs(uo - ur)) <= 1e-13 * np.max(np.abs(ur))
    else:
        assert np.array_equal(uo, ur)
    if ic[4]:
        assert np.array_equal(oracle.deriv()[:16], ref.deriv()[:16])


@pytest.mark.parametrize("algo,eqstate,coords", [(5, 1, 2), (5, 1, 1), (6, 1, 1), (4, 2, 1),
                                                 (1, 3, 1), (1, 3, 2), (2, 3, 1), (1, 1, 2), (2, 1, 1)])
@pytest.mark.parametrize("divclean", [0, 1])
def test_pointwise_mhd_bitexact(oracle, ref, algo, eqstate, coords, divclean):
    ic, rc = make_config(mhd=1, divclean=divclean, algo_solver=algo, eqstate=eqstate,
                         coordinates=coords)
    oracle.setup(ic, rc)
    ref.setup(ic, rc)
    t = 1.3 if coords == 2 else 1.0
    oracle.set_time(t)
    ref.set_time(t)
    nv = oracle.nv
    for _ in range(60):
        v = np.zeros(nv)
        v[0] = RNG.uniform(0.5, 2.0)
        v[1:4] = RNG.uniform(-0.5, 0.5, 3) / np.sqrt(3) * np.array([1, 1, 1 / t])
        v[4] = np.exp(RNG.uniform(np.log(0.05), np.log(20.0)))
        v[5:8] = RNG.uniform(-1, 1, 3) * np.sqrt(v[4])
        if divclean:
            v[8] = RNG.uniform(-0.1, 0.1)
        e1, u1, v1 = oracle.prim2cons(v)
        e2, u2, v2 = ref.prim2cons(v)
        assert e1 == e2 and np.array_equal(u1, u2)
        for d in (1, 2, 3):
            a = oracle.flux(v, d)
            b = ref.flux(v, d)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        guess = v.copy()
        guess[4] *= 1.05
        guess[1:4] *= 0.97
        a = oracle.cons2prim(u1, guess)
        b = ref.cons2prim(u1, guess)
        assert a[0] == b[0] and np.array_equal(a[1], b[1])
