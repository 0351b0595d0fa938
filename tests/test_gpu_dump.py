"""Restart dump on the GPU (echo_set_dump / echo_dump / echo_restart): the
dump.dat layout of out_dump with the oracle's values (itself byte-identical
to the reference's out_dump), a restart that continues bit for bit like the
uninterrupted run (also with the freeze-out finder active), and a restart
from a file the oracle wrote."""
import numpy as np
import pytest

from libs import OracleFO, oracle_dump, oracle_restart
from scenarios import fireball3d, mhd3d, rel_linf, rmhd_fireball

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _parse(path, nv, n, kend, fo=False):
    b = open(path, "rb").read()
    iout = int(np.frombuffer(b[:4], dtype=np.int32)[0])
    d = np.frombuffer(b[4:], dtype=np.float64)
    out = {"iout": iout, "t": d[0], "timeold": d[1], "timeinterval": d[2]}
    o = 3
    for name, m in (("v", nv * n), ("u", nv * n), ("vold", kend * n), ("uold", kend * n)):
        out[name] = d[o:o + m]
        o += m
    out["g3old"] = d[o]
    o += 1
    if fo:
        out["thyp"] = d[o]
        out["time_arr"] = d[o + 1:o + 3]
        out["vb"] = d[o + 3:o + 3 + nv * n]
        out["field2"] = d[o + 3 + nv * n:o + 3 + nv * n + n]
        o += 3 + nv * n + n
    assert o == d.size
    return out


def _grid(ic, rc):
    from echoamd import EchoGrid
    return EchoGrid(ic, rc)


@pytest.mark.parametrize("make,kend", [(lambda: fireball3d(24, 20, 16, viscosity=1, eqstate=3), 5),
                                       (lambda: mhd3d(16, 12, 10, algo=1), 8),
                                       (lambda: rmhd_fireball(16, 12, 10, viscosity=0, nrk=3), 8),
                                       (lambda: rmhd_fireball(16, 12, 10, viscosity=1), 15)],
                         ids=["visc-tab", "mhd-tab", "rmhd-rk3", "vrmhd-imex"])
def test_dump_matches_oracle(oracle, tmp_path, make, kend):
    print("\ndump vs oracle", flush=True)
    ic, rc, v0 = make()
    g = _grid(ic, rc)
    g.init_from_prim(v0)
    g.set_dump(True)
    g.steps(3)
    gp = tmp_path / "gpu_dump.dat"
    g.dump(gp, 6)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(3)
    op = tmp_path / "orc_dump.dat"
    assert oracle_dump(oracle, op, 6) == 0
    n = int(ic[0] * ic[1] * ic[2])
    nv = g.nv
    a, b = _parse(gp, nv, n, kend), _parse(op, nv, n, kend)
    assert a["iout"] == b["iout"] == 6
    for k in ("t", "timeold", "timeinterval", "g3old"):
        assert abs(a[k] - b[k]) <= 1e-13 * abs(b[k]), k
    for k, m in (("v", nv), ("u", nv), ("vold", kend), ("uold", kend)):
        assert rel_linf(a[k].reshape(m, -1), b[k].reshape(m, -1)) <= TOL, k
    # v and u of the file are the device state
    vs, us = g.state()
    assert np.array_equal(a["v"], vs.ravel()) and np.array_equal(a["u"], us.ravel())
    g.close()


@pytest.mark.parametrize("make", [lambda: fireball3d(24, 20, 16, viscosity=1, eqstate=3),
                                  lambda: rmhd_fireball(16, 12, 10, viscosity=1)], ids=["visc-tab", "vrmhd-imex"])
def test_restart_continues_bit_for_bit(tmp_path, make):
    print("\nrestart round trip", flush=True)
    ic, rc, v0 = make()
    g = _grid(ic, rc)
    g.init_from_prim(v0)
    g.set_dump(True)
    g.steps(3)
    path = tmp_path / "dump.dat"
    g.dump(path, 11)
    g.steps(4)
    va, ua = g.state()
    sa = g.scalars()
    g.close()
    h = _grid(ic, rc)
    h.set_dump(True)
    iout, _ = h.restart(path)
    assert iout == 11
    h.steps(4)
    vb, ub = h.state()
    sb = h.scalars()
    assert np.array_equal(va, vb) and np.array_equal(ua, ub)
    assert sa["t"] == sb["t"]
    h.close()


def test_restart_from_oracle_dump(oracle, tmp_path):
    print("\nrestart from the oracle's file", flush=True)
    ic, rc, v0 = fireball3d(24, 20, 16, viscosity=1, eqstate=3)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(3)
    path = tmp_path / "orc_dump.dat"
    assert oracle_dump(oracle, path, 2) == 0
    oracle.steps(3)
    vo, uo = oracle.state()
    g = _grid(ic, rc)
    iout, _ = g.restart(path)
    assert iout == 2
    g.steps(3)
    vg, ug = g.state()
    assert rel_linf(vg, vo) <= TOL and rel_linf(ug, uo) <= TOL
    g.close()


def test_restart_with_freezeout(tmp_path):
    print("\nrestart with the freeze-out finder", flush=True)
    ic, rc, v0 = fireball3d(24, 20, 16, viscosity=1, eqstate=3)
    g = _grid(ic, rc)
    g.init_from_prim(v0)
    g.set_dump(True)
    g.fo_setup(0, 0.2)
    assert g.fo_slice() == 0
    g.steps(2)
    g.fo_slice()
    path = tmp_path / "dump.dat"
    g.dump(path, 3, thyp=1.5)
    n = int(ic[0] * ic[1] * ic[2])
    d = _parse(path, g.nv, n, 5, fo=True)
    assert d["thyp"] == 1.5
    vs, _ = g.state()
    assert np.array_equal(d["vb"], vs.ravel())      # the slice the next comparison uses
    g.steps(2)
    na = g.fo_slice()
    ea = g.fo_elements()
    g.close()
    h = _grid(ic, rc)
    h.fo_setup(0, 0.2)
    h.set_dump(True)
    iout, thyp = h.restart(path)
    assert iout == 3 and thyp == 1.5
    h.steps(2)
    nb = h.fo_slice()
    eb = h.fo_elements()
    assert na == nb > 0 and np.array_equal(ea, eb)
    h.close()


@pytest.mark.parametrize("how", ["timed", "stages"])
def test_dump_after_timed_and_single_stage_steps(oracle, tmp_path, how):
    """vold / uold are tracked however the steps are taken: with kernel
    timing on (plain launches instead of the captured graph) and through
    echo_rk_stage / echo_end_step."""
    ic, rc, v0 = fireball3d(24, 20, 16, viscosity=1, eqstate=3)
    g = _grid(ic, rc)
    g.init_from_prim(v0)
    g.set_dump(True)
    if how == "timed":
        g.set_timing(True)
        g.steps(3)
    else:
        for _ in range(3):
            for irk in (1, 2, 3):
                g.rk_stage(irk)
            g.end_step()
    gp = tmp_path / "gpu_dump.dat"
    g.dump(gp, 6)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(3)
    op = tmp_path / "orc_dump.dat"
    assert oracle_dump(oracle, op, 6) == 0
    n = int(ic[0] * ic[1] * ic[2])
    a, b = _parse(gp, g.nv, n, 5), _parse(op, g.nv, n, 5)
    for k, m in (("v", g.nv), ("u", g.nv), ("vold", 5), ("uold", 5)):
        assert rel_linf(a[k].reshape(m, -1), b[k].reshape(m, -1)) <= TOL, k
    g.close()
