"""Restart dump: the oracle's restatement of out_dump / out_restart
(oracle_out.c) against the reference's own out_dump (oracle/_ref), byte for
byte, with and without the freeze-out section; and the round trip
dump -> restart -> steps equals the uninterrupted run bit for bit."""
import numpy as np
import pytest

from libs import OracleFO, RefFO, oracle_dump, oracle_restart, ref_dump
from scenarios import fireball3d, mhd3d

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
