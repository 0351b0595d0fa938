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
