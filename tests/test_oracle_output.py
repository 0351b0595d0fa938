"""Output frames: the oracle's restatement of out_vars (oracle_out.c)
against the reference's own out_vars (oracle/_ref), byte for byte, for
viscous and MHD states, FP64 and FP32 frames."""
import ctypes as C
import os

import numpy as np
import pytest

from libs import EOS_DIR
from scenarios import fireball3d, mhd3d, reference_resistive_config, rmhd_fireball

ALL = np.ones(28, dtype=np.int32)
ALL[27] = 0          # rc (comoving density) is not written


def oracle_frame(oracle, sel, prec, iout, path, tpath):
    f = oracle.lib.orc_out_frame
    f.restype = C.c_int
    f.argtypes = [np.ctypeslib.ndpointer(np.int32), C.c_int, C.c_int, C.c_char_p, C.c_char_p]
    return f(np.ascontiguousarray(sel, dtype=np.int32), prec, iout, path.encode(), tpath.encode())


def ref_frame(ref, sel, prec, iout, outdir):
    f = ref.lib.ref_out_vars
    f.restype = None
    f.argtypes = [np.ctypeslib.ndpointer(np.int32), C.c_int, C.c_int, C.c_int, C.c_char_p]
    os.makedirs(os.path.join(outdir, "outr0001"), exist_ok=True)
    cwd = os.getcwd()
    os.chdir(EOS_DIR)
    try:
        f(np.ascontiguousarray(sel, dtype=np.int32), prec, iout, len(outdir), outdir.encode())
    finally:
        os.chdir(cwd)
    return os.path.join(outdir, "outr0001", f"out{iout:04d}.dat")


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("name,make,steps", [
    ("visc-tab", lambda: fireball3d(12, 10, 8, viscosity=1, eqstate=3), 3),
    ("mhd-glm-bjorken", lambda: mhd3d(12, 10, 8, algo=1), 3),
])
def test_frame_bytes_match_reference(oracle, ref, tmp_path, prec, name, make, steps):
    ic, rc, v0 = make()
    for lib in (oracle, ref):
        lib.setup(ic, rc)
        assert lib.init_from_prim(v0) == 0
        assert lib.steps(steps) == 0
    path_r = ref_frame(ref, ALL, prec, 7, str(tmp_path / "ref"))
    path_o = str(tmp_path / "orc.dat")
    assert oracle_frame(oracle, ALL, prec, 7, path_o, str(tmp_path / "time.dat")) == 0
    a = open(path_r, "rb").read()
    b = open(path_o, "rb").read()
    assert len(a) == len(b) and a == b
    tr = open(tmp_path / "ref" / "outr0001" / "time.dat").read()
    to = open(tmp_path / "time.dat").read()
    assert tr == to


def test_frame_selection_subset(oracle, ref, tmp_path):
    ic, rc, v0 = fireball3d(12, 10, 8, viscosity=1, eqstate=3)
    for lib in (oracle, ref):
        lib.setup(ic, rc)
        lib.init_from_prim(v0)
        lib.steps(2)
    sel = np.zeros(28, dtype=np.int32)
    sel[[0, 4, 6, 10, 16, 19]] = 1      # rho, p, T, pi^tx, pi^xx, v0
    path_r = ref_frame(ref, sel, 8, 2, str(tmp_path / "ref"))
    path_o = str(tmp_path / "orc.dat")
    assert oracle_frame(oracle, sel, 8, 2, path_o, str(tmp_path / "time.dat")) == 0
    assert open(path_r, "rb").read() == open(path_o, "rb").read()


@pytest.mark.parametrize("prec", [8, 4])
def test_resistive_frame_bytes_match_reference(oracle, ref, tmp_path, prec):
    """A resistive grid (B, E, psi after the fluid variables): the evolved E
    is written as held (out.f90:404-413), not the ideal -v x B.  The oracle
    evolves the grid; the reference's writer gets the oracle's state."""
    ic, rc, v0 = rmhd_fireball(12, 10, 8, viscosity=0)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(2) == 0
    ref.setup(reference_resistive_config(ic), rc)
    assert ref.init_from_prim(v0) == 0
    v, u = oracle.state()
    ref.set_state(v, u)
    ref.set_time(oracle.scalars()["t"])
    path_r = ref_frame(ref, ALL, prec, 3, str(tmp_path / "ref"))
    path_o = str(tmp_path / "orc.dat")
    assert oracle_frame(oracle, ALL, prec, 3, path_o, str(tmp_path / "time.dat")) == 0
    a = open(path_r, "rb").read()
    b = open(path_o, "rb").read()
    assert len(a) == len(b) and a == b
    # the E fields of the frame are the state's E, which differs from -v x B
    n = v[0].size
    dt = np.float64 if prec == 8 else np.float32
    f = np.fromfile(path_o, dtype=dt)[1:].reshape(-1, n)
    assert np.array_equal(f[12:15], v[8:11].reshape(3, n).astype(dt))
    ideal_ex = -(v[2] * v[7] - v[3] * v[6]).ravel()
    assert np.abs(f[12] - ideal_ex).max() > 1e-3
