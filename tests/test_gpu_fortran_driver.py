"""The Fortran caller (echo-qgp-mhd-devel_amd/fortran/echo_amd_driver, built
with flang on the iso_c_binding module echo_amd_f) runs a param.dat-format
Gubser case on the GPU; its final state equals the oracle's to 1e-10."""
import os
import subprocess

import numpy as np
import pytest

from libs import PKG, Oracle
from scenarios import GUBSER_THAT0, gubser2d, rel_linf

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(PKG, "fortran", "echo_amd_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_gubser(oracle, tmp_path):
    n, steps = 32, 12
    ic, rc, v0 = gubser2d(n=n)
    half = n * 0.1 / 2
    dump = tmp_path / "state.bin"
    lines = ["! Gubser flow, param.dat format", "INIT_TYPE=4", "COORD....=2",
             f"NX.......={n}", f"NY.......={n}", "NZ.......=1",
             f"XMIN.....={-half!r}", f"XMAX.....={half!r}", f"YMIN.....={-half!r}",
             f"YMAX.....={half!r}", "ZMIN.....=-0.5", "ZMAX.....=0.5", "TSTART...=1.0",
             "CFL......=0.4", "MAXDT....=1.0", "REC_ALGO.=WENO5", "INT_ALGO.=RK3",
             "EOS......=1", "bound_con=2", f"GUB_T0...={GUBSER_THAT0!r}",
             f"NSTEPS...={steps}", "CHUNK....=5", f"DUMP_FILE={dump}"]
    pf = tmp_path / "param.dat"
    pf.write_text("\n".join(lines) + "\n")
    r = subprocess.run([DRIVER, str(pf)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:])
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(dump, dtype=np.float64)
    nv = 5
    v = raw[:nv * n * n].reshape(nv, 1, n, n)
    u = raw[nv * n * n:].reshape(nv, 1, n, n)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(steps) == 0
    vo, uo = oracle.state()
    assert rel_linf(v, vo) <= 1e-10
    assert rel_linf(u, uo) <= 1e-10


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_freezeout(oracle, tmp_path):
    """HYP_COMPU=1: the driver writes hypersurface.txt at the reference's
    cadence (echo.f90:252-258); the oracle, run through the same cadence,
    writes the same file to the 8 digits of its format."""
    from libs import OracleFO
    n, steps, fv, fdt = 32, 12, 0.7, 0.02
    ic, rc, v0 = gubser2d(n=n)
    half = n * 0.1 / 2
    hyp = tmp_path / "hyp_gpu.txt"
    lines = ["INIT_TYPE=4", "COORD....=2", f"NX.......={n}", f"NY.......={n}", "NZ.......=1",
             f"XMIN.....={-half!r}", f"XMAX.....={half!r}", f"YMIN.....={-half!r}",
             f"YMAX.....={half!r}", "ZMIN.....=-0.5", "ZMAX.....=0.5", "TSTART...=1.0",
             "CFL......=0.4", "MAXDT....=1.0", "REC_ALGO.=WENO5", "INT_ALGO.=RK3",
             "EOS......=1", "bound_con=2", f"GUB_T0...={GUBSER_THAT0!r}",
             f"NSTEPS...={steps}", "HYP_COMPU=1", "FREEZKIND=0", f"FREEZEVAL={fv!r}",
             f"HYPSURFTI={fdt!r}", f"HYP_FILE.={hyp}"]
    pf = tmp_path / "param.dat"
    pf.write_text("\n".join(lines) + "\n")
    r = subprocess.run([DRIVER, str(pf)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:])
    assert r.returncode == 0, r.stdout + r.stderr
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    fo = OracleFO(oracle)
    assert fo.setup(0, fv)[0] == 0
    opath = str(tmp_path / "hyp_orc.txt")
    assert fo.write(opath, 1) == 0
    thyp, tmax, first = rc[6], 1e6, True
    for _ in range(steps):
        assert oracle.steps(1) == 0
        t = oracle.scalars()["t"]
        if t > thyp or thyp - t < 1e-10:
            thyp = min(thyp + fdt, tmax)
            assert fo.slice()[0] == 0
            if not first:
                assert fo.write(opath, 0) == 0
            first = False
    a = open(hyp).read().splitlines()
    b = open(opath).read().splitlines()
    assert len(a) == len(b) > 100
    assert a[0].split()[:2] == b[0].split()[:2]
    for la, lb in zip(a[1:], b[1:]):
        va, vb = np.array(la.split(), dtype=float), np.array(lb.split(), dtype=float)
        assert np.all(np.abs(va - vb) <= 1.5e-7 * np.abs(vb) + 1e-300), (la, lb)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_restart(oracle, tmp_path):
    """RESTART..=1 writes dump.dat at the output times; RESUME...= starts a
    second run from it.  The oracle, restarted from the same file, takes the
    same steps to 1e-10."""
    from libs import oracle_restart
    n, dtout = 32, 0.05
    ic, rc, v0 = gubser2d(n=n)
    half = n * 0.1 / 2
    base = ["INIT_TYPE=4", "COORD....=2", f"NX.......={n}", f"NY.......={n}", "NZ.......=1",
            f"XMIN.....={-half!r}", f"XMAX.....={half!r}", f"YMIN.....={-half!r}",
            f"YMAX.....={half!r}", "ZMIN.....=-0.5", "ZMAX.....=0.5", "TSTART...=1.0",
            "CFL......=0.4", "MAXDT....=1.0", "REC_ALGO.=WENO5", "INT_ALGO.=RK3",
            "EOS......=1", "bound_con=2", f"GUB_T0...={GUBSER_THAT0!r}", f"DTOUT....={dtout!r}"]
    run1 = tmp_path / "run1.dat"
    run1.write_text("\n".join(base + ["NSTEPS...=6", "RESTART..=1", f"DUMP_DIR.={tmp_path}"]) + "\n")
    r = subprocess.run([DRIVER, str(run1)], cwd=PKG, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dump = tmp_path / "dump.dat"
    assert dump.exists()
    final = tmp_path / "final.bin"
    run2 = tmp_path / "run2.dat"
    run2.write_text("\n".join(base + ["NSTEPS...=5", f"RESUME...={dump}", f"DUMP_FILE={final}"]) + "\n")
    r = subprocess.run([DRIVER, str(run2)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "restarted from" in r.stdout
    raw = np.fromfile(final, dtype=np.float64)
    nv = 5
    v = raw[:nv * n * n].reshape(nv, 1, n, n)
    u = raw[nv * n * n:].reshape(nv, 1, n, n)
    rc2 = rc.copy()
    rc2[16] = dtout
    oracle.setup(ic, rc2)
    e, iout, _ = oracle_restart(oracle, dump)
    assert e == 0 and iout >= 0
    assert oracle.steps(5) == 0
    vo, uo = oracle.state()
    assert rel_linf(v, vo) <= 1e-10
    assert rel_linf(u, uo) <= 1e-10
