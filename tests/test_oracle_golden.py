"""The oracle against the golden fixtures that tests/golden/make_golden.py
wrote from the reference Fortran (bit for bit).  Runs without the reference
library, so the oracle stays pinned on machines that do not have it."""
import os

import numpy as np
import pytest

from libs import REC_NGC, make_config

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} missing")
    return np.load(path)


def test_golden_reconstruction(oracle):
    g = load("rec.npz")
    for code in REC_NGC:
        for row, gl, gr in zip(g[f"rec{code}_f"], g[f"rec{code}_gl"], g[f"rec{code}_gr"]):
            a, b = oracle.rec(code, row)
            assert np.array_equal(a, gl) and np.array_equal(b, gr), code


def test_golden_eos(oracle):
    g = load("eos.npz")
    rh, en, pr = g["eos_rh"], g["eos_en"], g["eos_pr"]
    for eq in (1, 2, 3):
        ic, rc = make_config(eqstate=eq)
        oracle.setup(ic, rc)
        assert oracle.c_przero() == g[f"eos{eq}_przero"]
        P = np.array([oracle.eos_pressure(a, b)[1] for a, b in zip(rh, en)])
        E = np.array([oracle.eos_energy(a, b)[1] for a, b in zip(rh, pr)])
        C = np.array([oracle.eos_sound(a, b, c)[1] for a, b, c in zip(rh, en, pr)])
        D = np.array([oracle.eos_derived(a, b)[1] for a, b in zip(rh, pr)])
        assert np.array_equal(P, g[f"eos{eq}_pressure"])
        assert np.array_equal(E, g[f"eos{eq}_energy"])
        assert np.array_equal(C, g[f"eos{eq}_sound"])
        assert np.array_equal(D, g[f"eos{eq}_derived"])


@pytest.mark.parametrize("visc", [0, 1])
def test_golden_system(oracle, visc):
    g = load("system.npz")
    ic, rc = make_config(viscosity=visc, eqstate=3, coordinates=2)
    oracle.setup(ic, rc)
    oracle.set_time(0.6)
    for i, row in enumerate(g[f"sys{visc}_v"]):
        e, u, vout = oracle.prim2cons(row)
        assert np.array_equal(u, g[f"sys{visc}_u"][i])
        assert np.array_equal(vout, g[f"sys{visc}_vout"][i])
        for d in (1, 2, 3):
            _, _, f, vf = oracle.flux(row, d)
            assert np.array_equal(f, g[f"sys{visc}_f"][i, d - 1])
            assert np.array_equal(vf, g[f"sys{visc}_vf"][i, d - 1])
        guess = row.copy()
        guess[4] *= 1.1
        assert np.array_equal(oracle.cons2prim(u, guess)[1], g[f"sys{visc}_c2p"][i])


@pytest.mark.parametrize("name", ["gubser16_weno5", "shock16_mpe5", "fireball_visc_tab",
                                  "fireball_ideal_tab_rk2_mink"])
def test_golden_trajectory(oracle, name):
    g = load(f"traj_{name}.npz")
    oracle.setup(g["ic"], g["rc"])
    assert oracle.init_from_prim(g["v0"]) == 0
    assert oracle.steps(int(g["nsteps"])) == 0
    v, u = oracle.state()
    s = oracle.scalars()
    assert s["t"] == g["t"] and s["dt"] == g["dt"]
    assert np.array_equal(v, g["v"]) and np.array_equal(u, g["u"])


REPAIR_CASES = ["rough-ideal-weno5", "rough-ideal-mpe5", "rough-table-weno5", "forced-idealgas-tvd2",
                "forced-idealgas-weno5", "forced-viscous-weno5", "fast-eta-crash"]


@pytest.mark.parametrize("name", REPAIR_CASES)
def test_golden_repairs(oracle, name):
    """the repair branches (evolve.f90:334-369, 795-842) on the states of
    tests/test_oracle_repair.py, against the reference's results"""
    g = load("repair.npz")
    oracle.setup(g[f"{name}/ic"], g[f"{name}/rc"])
    assert oracle.init_from_prim(g[f"{name}/v0"]) == 0
    if f"{name}/v_in" in g:
        oracle.set_state(g[f"{name}/v_in"], g[f"{name}/u_in"])
    assert oracle.steps(int(g[f"{name}/steps"])) == int(g[f"{name}/err"])
    cnt = oracle.repair_counts()
    assert cnt["u2v"] > 0 and cnt["c2p_undef"] == 0
    v, u = oracle.state()
    assert np.array_equal(v, g[f"{name}/v"], equal_nan=True)
    assert np.array_equal(u, g[f"{name}/u"], equal_nan=True)


def test_golden_u2v_alone(oracle):
    g = load("repair.npz")
    oracle.setup(g["u2v/ic"], g["u2v/rc"])
    oracle.init_from_prim(g["u2v/v0"])
    oracle.set_state(g["u2v/v_in"], g["u2v/u_in"])
    assert oracle.u2v() == int(g["u2v/err"])
    assert oracle.repair_counts()["u2v_pr"] > 0
    v, u = oracle.state()
    assert np.array_equal(v, g["u2v/v"]) and np.array_equal(u, g["u2v/u"])
