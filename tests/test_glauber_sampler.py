"""The Glauber-MC event generator (echo_glauber_generate, host code of the
C-ABI) and its event files, on the CPU.

Its random stream is its own (the reference's gfortran rand() is not
reproduced), so it is pinned by what the sampling fixes:
- the collision test and the files: from the participants and spectators
  of an event the collisions are recomputed by brute force (d^2 <=
  sigma_NN / pi, sigma in fm^2 = mb / 10) and must match the written
  collisions and the participant / spectator split exactly; the first Z
  nucleons of each nucleus are protons;
- the nucleon positions: transverse radii against an independent numpy
  sampler of the Woods-Saxon density (two-sample Kolmogorov-Smirnov);
- the impact parameter: p(b) ~ b (KS against b^2 / 100 on b < 10 fm, where
  every configuration gives an event);
- published Glauber numbers at b = 0: Au-Au at 42 mb and Pb-Pb at 67.6 mb,
  mean N_part and N_coll within the ranges of the literature;
- the binary layout of the files (20-byte and 16-byte records).
"""
import numpy as np
import pytest
from scipy import stats

import libs  # noqa: F401  (puts the package on the path)
from echoamd import NUCLEI, EchoError, glauber_generate, glauber_read_event


def _events(tmp_path, n_max=None, **kw):
    n = glauber_generate(tmp_path, **kw)
    m = n if n_max is None else min(n, n_max)
    return n, [glauber_read_event(tmp_path, i + 1) for i in range(m)]


def test_collisions_and_participants_exact(tmp_path):
    n, evs = _events(tmp_path, seed=3, nucleus="Au", nconf=6, nbcoll=4, n_max=10)
    assert n > 0
    s = 42.0 / 10.0 / np.pi
    for ev in evs:
        # the listed positions carry the +-b/2 shifts: their distance is the
        # (x_A - x_B + b, y_A - y_B) of the test, their midpoint the
        # collision point
        a = np.concatenate([ev["part_a"], ev["spec_a"]])
        bb = np.concatenate([ev["part_b"], ev["spec_b"]])
        assert len(a) == len(bb) == 197
        d2 = ((a[:, None, :] - bb[None, :, :]) ** 2).sum(-1)
        hit = d2 <= s
        na, nb = len(ev["part_a"]), len(ev["part_b"])
        assert hit[:na].any(1).all() and not hit[na:].any()
        assert hit[:, :nb].any(0).all() and not hit[:, nb:].any()
        ia, ib = np.nonzero(hit)
        mid = (a[ia] + bb[ib]) / 2
        assert len(ev["coll"]) == hit.sum()
        got = ev["coll"][np.lexsort(ev["coll"].T[::-1])]
        want = mid[np.lexsort(mid.T[::-1])]
        assert np.allclose(got, want, rtol=0, atol=1e-12)
        assert len(ev["coll"]) >= 1 and na + nb >= 18
        # protons: 79 of each nucleus
        assert (ev["type_part_a"] == 0).sum() + (ev["type_spec_a"] == 0).sum() == 79
        assert (ev["type_part_b"] == 0).sum() + (ev["type_spec_b"] == 0).sum() == 79


def _ws_transverse(rng, nuc, n_nuclei):
    """independent sampler: inverse-CDF Woods-Saxon radii, isotropic
    directions, each nucleus moved to its transverse centre of mass"""
    r = np.linspace(0, 12.6, 20001)
    pdf = r ** 2 / (np.exp((r - nuc["radius"]) / nuc["diffuse"]) + 1)
    cdf = np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) / 2)])
    cdf /= cdf[-1]
    out = []
    for _ in range(n_nuclei):
        rr = np.interp(rng.uniform(size=nuc["A"]), cdf, r)
        ct = rng.uniform(-1, 1, nuc["A"])
        ph = rng.uniform(0, 2 * np.pi, nuc["A"])
        st = np.sqrt(1 - ct ** 2)
        xy = np.stack([rr * st * np.cos(ph), rr * st * np.sin(ph)], 1)
        out.append(xy - xy.mean(0))
    return np.concatenate(out)


def test_woods_saxon_positions(tmp_path):
    n, evs = _events(tmp_path, seed=5, nucleus="Pb", nconf=60, nbcoll=1, b=0.0)
    assert n == 60
    rt = np.concatenate([np.hypot(*np.concatenate([e["part_a"], e["spec_a"]]).T) for e in evs] +
                        [np.hypot(*np.concatenate([e["part_b"], e["spec_b"]]).T) for e in evs])
    ref = np.hypot(*_ws_transverse(np.random.default_rng(1), NUCLEI["Pb"], 120).T)
    assert stats.ks_2samp(rt, ref).pvalue > 1e-3
    # a wrong diffuseness is seen
    wrong = np.hypot(*_ws_transverse(np.random.default_rng(1), dict(NUCLEI["Pb"], diffuse=0.42), 120).T)
    assert stats.ks_2samp(rt, wrong).pvalue < 1e-4


def test_impact_parameter_distribution(tmp_path):
    n, evs = _events(tmp_path, seed=7, nucleus="Au", nconf=1000, nbcoll=3)
    b = np.array([e["b"] for e in evs])
    assert b.max() <= 20.0 and b.min() >= 0.0
    small = b[b < 10.0]
    assert len(small) > 200
    assert stats.kstest(small, lambda x: (x / 10.0) ** 2).pvalue > 1e-3


@pytest.mark.parametrize("nucleus,sigma,npart,ncoll", [
    ("Au", 42.0, (355, 394), (1000, 1400)),
    ("Pb", 67.6, (390, 416), (1700, 2400)),
])
def test_central_collision_numbers(tmp_path, nucleus, sigma, npart, ncoll):
    """b = 0: <N_part> and <N_coll> within the published ranges (Au-Au at
    200 GeV: N_part ~ 370-380, N_coll ~ 1100-1250 at b = 0; Pb-Pb at 5 TeV:
    N_part ~ 405, N_coll ~ 1900-2100 at b = 0)."""
    n, evs = _events(tmp_path, seed=11, nucleus=nucleus, sigma_nn=sigma, nconf=40, nbcoll=1, b=0.0)
    assert n == 40
    npm = np.mean([len(e["part_a"]) + len(e["part_b"]) for e in evs])
    ncm = np.mean([len(e["coll"]) for e in evs])
    assert npart[0] <= npm <= npart[1], npm
    assert ncoll[0] <= ncm <= ncoll[1], ncm


def test_file_layout_and_determinism(tmp_path):
    d1, d2, d3 = tmp_path / "a", tmp_path / "b", tmp_path / "c"
    for d in (d1, d2, d3):
        d.mkdir()
    n1 = glauber_generate(d1, seed=21, nconf=5, nbcoll=2)
    n2 = glauber_generate(d2, seed=21, nconf=5, nbcoll=2)
    glauber_generate(d3, seed=22, nconf=5, nbcoll=2)
    for f in ("summary.dat", "partcoll.bin", "spectators.bin", "collisions.bin"):
        assert (d1 / f).read_bytes() == (d2 / f).read_bytes()
    assert (d1 / "partcoll.bin").read_bytes() != (d3 / "partcoll.bin").read_bytes()
    rec = np.dtype([("x", "<f8"), ("y", "<f8"), ("t", "<i4")])
    part = np.fromfile(d1 / "partcoll.bin", dtype=rec)
    spec = np.fromfile(d1 / "spectators.bin", dtype=rec)
    coll = np.fromfile(d1 / "collisions.bin", dtype="<f8").reshape(-1, 2)
    evs = [glauber_read_event(d1, i + 1) for i in range(n1)]
    assert len(part) == sum(len(e["part_a"]) + len(e["part_b"]) for e in evs)
    assert len(spec) == 2 * 197 * n1 - len(part)
    assert len(coll) == sum(len(e["coll"]) for e in evs)
    last = evs[-1]
    k = len(last["part_a"]) + len(last["part_b"])
    assert np.array_equal(np.stack([part["x"][-k:], part["y"][-k:]], 1),
                          np.concatenate([last["part_a"], last["part_b"]]))
    assert np.array_equal(part["t"][-k:], np.concatenate([last["type_part_a"], last["type_part_b"]]))
    assert n1 == n2


def test_rejects_bad_input(tmp_path):
    with pytest.raises(EchoError, match="A <= 300"):
        glauber_generate(tmp_path, seed=1, nucleus=dict(NUCLEI["Au"], A=400))
    with pytest.raises(EchoError, match="cannot create"):
        glauber_generate(tmp_path / "missing", seed=1)
    glauber_generate(tmp_path, seed=1, nconf=2)
    with pytest.raises(EchoError, match="no such event"):
        glauber_read_event(tmp_path, 10 ** 6)
