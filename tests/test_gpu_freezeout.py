"""GPU freeze-out hypersurface finder (echo_fo_*) against the oracle's
restatement of out_hyp / find_hypersurface_* (itself bit-exact against the
reference): the oracle is fed the GPU's own slices (state and time), so
both finders see the same input; the elements must come in the same order,
equal to 1e-12 (the GPU build contracts the interpolations to FMA), and the
text file must carry the same values.  At a large size the elements are
checked through properties the geometry fixes: with p = e/3 and freeze-out
on e every element has p = fv/3, the spatial normals of a closed surface
cancel, and the element count equals a direct count of the crossings."""
import numpy as np
import pytest

from libs import OracleFO
from scenarios import fireball3d, mhd3d, rmhd_fireball
from test_oracle_freezeout import SCENARIOS

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _feed(oracle, g):
    v, u = g.state()
    oracle.set_state(v, u)
    oracle.set_time(g.scalars()["t"])


def _close(a, b):
    """element rows equal to TOL relative to each column's scale"""
    assert a.shape == b.shape
    if a.size == 0:
        return
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    assert np.all(np.abs(a - b) <= TOL * scale), np.abs(a - b).max(axis=0) / scale


def _rows(path):
    """numeric rows of hypersurface.txt after the header, count lines dropped"""
    out = []
    for ln in open(path).read().splitlines()[1:]:
        if len(ln) == 6:
            continue
        out.append([float(x) for x in ln.split()])
    return np.array(out)


# resistive grids (find_hypersurface_mhd with B at the resistive slots)
RESISTIVE = [
    ("rmhd-rk3-T", lambda: rmhd_fireball(12, 10, 8, viscosity=0, nrk=3), 0, 0.2),
    ("vrmhd-imex-e", lambda: rmhd_fireball(12, 10, 8, viscosity=1), 1, 1.5),
]


@pytest.mark.parametrize("name,make,ftype,fval", SCENARIOS + RESISTIVE, ids=[s[0] for s in SCENARIOS + RESISTIVE])
def test_gpu_elements_match_oracle(oracle, tmp_path, name, make, ftype, fval):
    from echoamd import EchoGrid
    print(f"\nfreeze-out {name}", flush=True)
    ic, rc, v0 = make()
    g = EchoGrid(ic, rc)
    assert g.init_from_prim(v0) == 0
    oracle.setup(ic, rc)
    ofo = OracleFO(oracle)
    hdr = g.fo_setup(ftype, fval)
    e, ohdr = ofo.setup(ftype, fval)
    assert e == 0
    # the header values (T_f, e_f) and line
    assert np.allclose(hdr, ohdr, rtol=1e-12, atol=0)
    gpath, opath = str(tmp_path / "gpu.txt"), str(tmp_path / "orc.txt")
    g.fo_write(gpath, header=True)
    assert ofo.write(opath, 1) == 0
    hg, ho = open(gpath).read().split(), open(opath).read().split()
    assert len(hg) == len(ho) == 3 and hg[0] == ho[0] == str(ftype)
    assert np.allclose([float(x) for x in hg[1:]], [float(x) for x in ho[1:]], rtol=1e-12, atol=0)
    total = 0
    for k in range(3):
        if k:
            g.steps(2)
        _feed(oracle, g)
        n = g.fo_slice()
        e, no = ofo.slice()
        assert e == 0 and n == no, (n, no)
        if k == 0:
            assert n == 0
            continue
        _close(g.fo_elements(), ofo.elements())
        g.fo_write(gpath)
        assert ofo.write(opath, 0) == 0
        total += n
    assert total > 40
    a, b = _rows(gpath), _rows(opath)
    assert a.shape == b.shape
    assert np.all(np.abs(a - b) <= 1.5e-7 * np.maximum(np.abs(b), 1e-300) + 1e-300)
    g.close()


def test_gpu_large_closed_surface():
    """192 x 192 x 64 fireball, p = e/3, freeze-out on e: properties only."""
    from echoamd import EchoGrid
    print("\nfreeze-out large fireball 192x192x64", flush=True)
    ic, rc, v0 = fireball3d(192, 192, 64, eqstate=1)
    g = EchoGrid(ic, rc)
    assert g.init_from_prim(v0) == 0
    fv = 3.0
    g.fo_setup(1, fv)
    assert g.fo_slice() == 0
    v_old, _ = g.state()
    g.steps(3)
    v_new, _ = g.state()
    n = g.fo_slice()
    a = g.fo_elements()
    assert len(a) == n > 10000
    # direct count of the crossings on the two slices (e = 3 p)
    e1, e2 = 3.0 * v_old[4], 3.0 * v_new[4]

    def cross(x, y):
        return (y - fv) * (fv - x) > 0.0
    cnt = cross(e1, e2).sum() + cross(e1[:, :, :-1], e1[:, :, 1:]).sum() + \
        cross(e1[:, :-1, :], e1[:, 1:, :]).sum() + cross(e1[:-1], e1[1:]).sum()
    assert n == cnt
    assert np.allclose(a[:, 8], fv / 3, rtol=1e-13, atol=0)
    for k in (10, 11, 12):
        col = a[:, k]
        assert abs(col.sum()) <= 1e-12 * np.abs(col).sum()
    # reference order: ix slowest, then iy, then iz, tau before the spatial
    # normals; cells are recovered from the positions
    hx, hy, hz = 6.0 / 192, 6.0 / 192, 2.0 / 64
    d = np.argmax(a[:, 9:13] != 0, axis=1)
    ux, uy, uz = (a[:, 1] + 3.0) / hx - 0.5, (a[:, 2] + 3.0) / hy - 0.5, (a[:, 3] + 1.0) / hz - 0.5
    ix = np.where(d == 1, np.floor(ux), np.round(ux)).astype(np.int64)
    iy = np.where(d == 2, np.floor(uy), np.round(uy)).astype(np.int64)
    iz = np.where(d == 3, np.floor(uz), np.round(uz)).astype(np.int64)
    key = ((ix * 192 + iy) * 64 + iz) * 4 + d
    assert np.all(np.diff(key) > 0)
    g.close()
