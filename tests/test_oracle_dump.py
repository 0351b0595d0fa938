"""Restart dump: the oracle's restatement of out_dump / out_restart
(oracle_out.c) against the reference's own out_dump (oracle/_ref), byte for
byte, with and without the freeze-out section; and the round trip
dump -> restart -> steps equals the uninterrupted run bit for bit."""
import numpy as np
import pytest

from libs import OracleFO, RefFO, oracle_dump, oracle_restart, ref_dump
from scenarios import fireball3d, mhd3d, reference_resistive_config, rmhd_fireball

CASES = [
    ("visc-tab", lambda: fireball3d(12, 10, 8, viscosity=1, eqstate=3)),
    ("mhd-glm-tab", lambda: mhd3d(12, 10, 8, algo=1)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_dump_bytes_match_reference(oracle, ref, tmp_path, name, make):
    ic, rc, v0 = make()
    for lib in (oracle, ref):
        lib.setup(ic, rc)
        assert lib.init_from_prim(v0) == 0
        assert lib.steps(3) == 0
    rpath = ref_dump(ref, str(tmp_path / "ref"), 7)
    opath = tmp_path / "orc_dump.dat"
    assert oracle_dump(oracle, opath, 7) == 0
    a, b = open(rpath, "rb").read(), open(opath, "rb").read()
    assert len(a) == len(b) and a == b


def test_dump_with_freezeout_matches_reference(oracle, ref, tmp_path):
    ic, rc, v0 = fireball3d(12, 10, 8, viscosity=1, eqstate=3)
    for lib in (oracle, ref):
        lib.setup(ic, rc)
        assert lib.init_from_prim(v0) == 0
    ofo, rfo = OracleFO(oracle), RefFO(ref, str(tmp_path / "ref"))
    assert ofo.setup(0, 0.2)[0] == 0
    rfo.setup(0, 0.2)
    for k in range(2):
        if k:
            assert oracle.steps(2) == 0 and ref.steps(2) == 0
        assert ofo.slice()[0] == 0 and rfo.slice() == 0
    rpath = ref_dump(ref, str(tmp_path / "ref"), 4, thyp=1.25, setup=False)
    opath = tmp_path / "orc_dump.dat"
    assert oracle_dump(oracle, opath, 4, thyp=1.25) == 0
    a, b = open(rpath, "rb").read(), open(opath, "rb").read()
    assert len(a) == len(b) and a == b


def test_restart_continues_bit_for_bit(oracle, tmp_path):
    ic, rc, v0 = fireball3d(16, 12, 10, viscosity=1, eqstate=3)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(3) == 0
    path = tmp_path / "dump.dat"
    assert oracle_dump(oracle, path, 9) == 0
    assert oracle.steps(4) == 0
    va, ua = oracle.state()
    sa = oracle.scalars()
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    e, iout, _ = oracle_restart(oracle, path)
    assert e == 0 and iout == 9
    assert oracle.steps(4) == 0
    vb, ub = oracle.state()
    sb = oracle.scalars()
    assert np.array_equal(va, vb) and np.array_equal(ua, ub)
    assert sa["t"] == sb["t"]


@pytest.mark.parametrize("viscosity", [0, 1])
def test_resistive_dump(oracle, ref, tmp_path, viscosity):
    """Resistive grids: vold / uold hold krh..kbz (common.f90:379-386), 8
    variables, or 15 on the viscous resistive grid, whose B follows the
    viscous variables.  Without viscosity the dump of the initial state is
    byte-identical to the reference's out_dump given the same state (the
    reference does not evolve the resistive system); with and without it the
    oracle's dump -> restart -> steps continues the run bit for bit."""
    ic, rc, v0 = rmhd_fireball(12, 10, 8, viscosity=viscosity)
    n = v0[0].size
    if not viscosity:
        oracle.setup(ic, rc)
        assert oracle.init_from_prim(v0) == 0
        ref.setup(reference_resistive_config(ic), rc)
        assert ref.init_from_prim(v0) == 0
        v, u = oracle.state()
        ref.set_state(v, u)
        rpath = ref_dump(ref, str(tmp_path / "ref"), 2)
        opath = tmp_path / "orc0.dat"
        assert oracle_dump(oracle, opath, 2) == 0
        a, b = open(rpath, "rb").read(), open(opath, "rb").read()
        assert len(a) == len(b) and a == b
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(2) == 0
    path = tmp_path / "dump.dat"
    assert oracle_dump(oracle, path, 5) == 0
    nv = 19 if viscosity else 12
    kend = 15 if viscosity else 8
    assert path.stat().st_size == 4 + 8 * 3 + 8 * n * (2 * nv + 2 * kend) + 8
    assert oracle.steps(3) == 0
    va, ua = oracle.state()
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    e, iout, _ = oracle_restart(oracle, path)
    assert e == 0 and iout == 5
    assert oracle.steps(3) == 0
    vb, ub = oracle.state()
    assert np.array_equal(va, vb) and np.array_equal(ua, ub)
