"""Output frames written by the GPU (echo_write_frame) against the oracle's
restatement of out_vars (itself byte-identical to the reference writer):
same layout and values; the primitive arrays of the frame equal the device
state bit for bit; the frame is a snapshot taken when the call is made even
if the grid keeps stepping while it is written."""
import os

import numpy as np
import pytest

from scenarios import fireball3d, mhd3d, rmhd_fireball
from test_oracle_output import oracle_frame

pytestmark = pytest.mark.gpu


def _fields(ic):
    from echoamd import OUT_FIELDS
    f = [x for x in OUT_FIELDS if x != "rc"]
    if not ic[4]:
        f = [x for x in f if not x.startswith("pi") and x != "bulk"]
    if not ic[6]:
        f = [x for x in f if x not in ("bx", "by", "bz", "ex", "ey", "ez", "glm")]
    return f


def _rel(a, b):
    """max |a - b| relative to max |b| of one field"""
    a, b = a.astype(np.float64), b.astype(np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _read(path, n, prec):
    dt = np.float64 if prec == 8 else np.float32
    a = np.fromfile(path, dtype=dt)
    return a[0], a[1:].reshape(-1, n)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("make", [lambda: fireball3d(16, 12, 10, viscosity=1, eqstate=3),
                                  lambda: mhd3d(16, 12, 10, algo=1),
                                  lambda: rmhd_fireball(16, 12, 10, viscosity=0, nrk=3),
                                  lambda: rmhd_fireball(16, 12, 10, viscosity=1)],
                         ids=["visc-tab", "mhd-glm", "rmhd-rk3", "vrmhd-imex"])
def test_gpu_frame(oracle, tmp_path, make, prec):
    from echoamd import OUT_FIELDS, EchoGrid
    ic, rc, v0 = make()
    g = EchoGrid(ic, rc)
    g.init_from_prim(v0)
    g.steps(3)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(3)
    fields = _fields(ic)
    vstate, _ = g.state()
    g.write_frame(fields, 5, str(tmp_path), prec)
    g.steps(2)                            # keeps stepping while the frame is written
    g.wait_frames()
    sel = np.zeros(28, dtype=np.int32)
    for f in fields:
        sel[OUT_FIELDS.index(f)] = 1
    assert oracle_frame(oracle, sel, prec, 5, str(tmp_path / "orc.dat"),
                        str(tmp_path / "orc_time.dat")) == 0
    n = int(ic[0] * ic[1] * ic[2])
    tg, ag = _read(tmp_path / "out0005.dat", n, prec)
    to, ao = _read(tmp_path / "orc.dat", n, prec)
    assert ag.shape == ao.shape == (len(fields), n)
    assert abs(tg - to) <= 1e-12 * abs(to)
    tol = 1e-10 if prec == 8 else 1e-6
    for k, f in enumerate(fields):
        assert _rel(ag[k], ao[k]) <= tol, f
    # primitive arrays of the frame are the device state at the call
    prim = {"density": 0, "vx": 1, "vy": 2, "vz": 3, "pressure": 4}
    for f, k in prim.items():
        want = vstate[k].ravel()
        got = ag[fields.index(f)]
        assert np.array_equal(got, want if prec == 8 else want.astype(np.float32)), f
    line = open(tmp_path / "time.dat").read()
    assert line.split()[0] == "5"
    g.close()
