"""The grid update at the large single-event sizes (384 x 384 x 256 and
192 x 192 x 128 cells; viscous hydro with RK3, and viscous resistive MHD
with IMEX; tabulated EoS, WENO5 + HLL, Milne coordinates), where the oracle
is too slow for the whole grid: through a size-independent property (a
state symmetric under the reflection (x, eta) -> (-x, -eta) stays
symmetric), and against the oracle on the box around a failing cell.

The initial state is the benchmark's (a Glauber-MC energy density over a
floor of 0.05 GeV/fm^3, Navier-Stokes stresses), made exactly
reflection-even as (a + mirror(a)) / 2.  Under the reflection rho, p, v^y, Pi, pi^{xz}, pi^{xx},
pi^{yy} and pi^{zz} are even, and v^x, v^eta, pi^{xy} and pi^{yz} are odd.
A wrong index, sign or stencil offset in any sweep, source or con2prim breaks
the symmetry at O(1).  The build's own rounding is not exactly symmetric
(cell centres, the order of the sums), and the near-vacuum edges amplify
it: after three steps the asymmetry is 4.8e-12 of the scale, so the bound is
1e-10 of each variable's (or vector's) scale."""
import numpy as np
import pytest

from scenarios import GLAUBER_PAR, glauber_event, pbpb_config

pytestmark = pytest.mark.gpu

# parity of the 12 viscous components under (x, eta) -> (-x, -eta):
# rho vx vy veta p Pi pixy pixz piyz pixx piyy pizz
PARITY = np.array([1, -1, 1, -1, 1, 1, -1, 1, -1, 1, 1, 1], dtype=np.float64)
GROUPS = [[0], [1, 2, 3], [4], [5], [6, 7, 8, 9, 10, 11]]


def test_reflection_symmetry_384x384x256():
    from libs import make_config
    from echoamd import EchoGrid, read_eos_rows
    nx, ny, nz, dx, deta = 384, 384, 256, 0.08, 0.05
    ic, rc = make_config(nx=nx, ny=ny, nz=nz, coordinates=2, viscosity=1, bulk=0, eqstate=3,
                         rec="WENO5", nrk=3, bc=2, obtained_zz=1, init_type=3,
                         xlim=(-nx * dx / 2, nx * dx / 2), ylim=(-ny * dx / 2, ny * dx / 2),
                         zlim=(-nz * deta / 2, nz * deta / 2), tmin=0.4, cfl=0.4, maxdt=0.1,
                         eta_over_s=0.12, tau_pi_coeff=5.0, enezero=0.001)
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
    import bench
    eos = read_eos_rows()
    v0, _ = bench.workload_state(ic, rc, eos, seed=7, b=5.0)
    v0 = 0.5 * (v0 + v0[:, ::-1, :, ::-1] * PARITY[:, None, None, None])
    print(f"\nfull size {nx}x{ny}x{nz}: symmetric initial state and 3 steps", flush=True)
    g = EchoGrid(ic, rc, eos_rows=eos)
    assert g.init_from_prim(v0) == 0
    del v0
    g.steps(3)
    s = g.scalars()
    assert s["crashed"] == 0 and s["iter"] == 3
    v, _ = g.state()
    g.close()
    assert np.all(np.isfinite(v))
    mirrored = v[:, ::-1, :, ::-1] * PARITY[:, None, None, None]
    worst = 0.0
    for grp in GROUPS:
        scale = max(np.abs(v[k]).max() for k in grp)
        if scale == 0.0:   # Pi without bulk viscosity
            assert all(np.array_equal(v[k], mirrored[k]) for k in grp)
            continue
        for k in grp:
            worst = max(worst, float(np.abs(v[k] - mirrored[k]).max() / scale))
    # the odd fields carry flow: the check is not passed by zeros
    assert np.abs(v[1]).max() > 1e-3 and np.abs(v[3]).max() > 1e-4
    print(f"max reflection asymmetry / scale: {worst:.2e}", flush=True)
    assert worst <= 1e-10


# ---------------------------------------------------------------------------
# The unfloored Glauber-MC state of the large grid (echo_init_glauber,
# ENEZERO 0.001).  Near the beam-rapidity cut of the deposition (|eta| ->
# ybeam = 5.36) the energy density falls to the vacuum value, the
# Navier-Stokes stresses exceed the pressure and v^eta approaches 1.  The
# first step needs no con2prim repair; in the second, thousands of cells are
# repaired and one repair fails, which stops the run (evolve.f90:836-840).
# The oracle (which matches the reference bit for bit on the repair paths)
# shows the same on the cells around the failure.
# ---------------------------------------------------------------------------
def _glauber_state(ic, rc, pa, pb, pc):
    from libs import Oracle, oracle_glauber_prim
    o = Oracle()
    o.setup(ic, rc)
    e, v0 = oracle_glauber_prim(o, GLAUBER_PAR, pa, pb, pc)
    assert e == 0
    return o, v0


def test_glauber_rapidity_tail_slab_vs_oracle():
    """An eta slab of the large grid around eta = 5.2 (384 x 384 x 16 cells,
    outflow boundaries of its own): step 1 with its repairs against the
    oracle, then the failed repair of step 2 with the same error code, the
    oracle's crashing cell in the message, and the cells before it (in the
    reference's loop order) equal."""
    from echoamd import EchoError, EchoGrid, read_eos_rows
    from scenarios import rel_linf
    ic, rc = pbpb_config(nz=16, zc=5.2)
    pa, pb, pc = glauber_event(seed=5)
    o, v0 = _glauber_state(ic, rc, pa, pb, pc)
    assert o.init_from_prim(v0) == 0
    g = EchoGrid(ic, rc, eos_rows=read_eos_rows())
    assert g.init_glauber(GLAUBER_PAR, pa, pb, pc) == 0
    print("\nslab 384x384x16 at eta 5.2: step 1", flush=True)
    assert o.steps(1) == 0
    g.steps(1)
    n1 = o.repair_counts()["u2v"]
    assert n1 > 0 and g.repairs() == n1
    vg, ug = g.state()
    vo, uo = o.state()
    assert rel_linf(vg, vo) <= 1e-10 and rel_linf(ug, uo) <= 1e-10
    print("step 2", flush=True)
    assert o.steps(1) == 1
    cnt = o.repair_counts()
    assert cnt["u2v_fail"] == 1
    crash = o.crash_cell()
    with pytest.raises(EchoError) as ei:
        g.steps(1)
    assert ei.value.code == 1 and "repair failed" in str(ei.value)
    print(str(ei.value), flush=True)
    if crash >= 0:
        nx, ny = int(ic[0]), int(ic[1])
        assert f"ix={crash % nx + 1} iy={(crash // nx) % ny + 1} iz={crash // (nx * ny) + 1}" in str(ei.value)
        vg, _ = g.state()
        vo, _ = o.state()
        before = np.arange(vg[0].size) < crash
        a = vg.reshape(12, -1)[:, before].reshape(12, 1, 1, -1)
        b = vo.reshape(12, -1)[:, before].reshape(12, 1, 1, -1)
        assert rel_linf(a, b) <= 1e-10
    g.close()


def test_glauber_unfloored_384x384x256_vs_oracle_box():
    """The whole grid from echo_init_glauber's unfloored state on the GPU:
    step 1 runs without repairs, step 2 stops on a failed con2prim repair,
    reported with its cell.  The oracle runs the 52^3-cell box around that
    cell (outflow boundaries; a face on the grid's own boundary is the same
    boundary) from its own Glauber state, with the GPU's time steps (its own
    CFL step is not smaller: the box holds a subset of the grid's speeds).
    After step 1 the box's interior (9 cells = 3 stages x 3 stencil cells
    away from the box's inner faces) equals the GPU's grid there; in step 2
    the oracle's run fails at the same cell."""
    import re
    from echoamd import EchoError, EchoGrid, read_eos_rows
    from libs import make_config
    from scenarios import rel_linf
    ic, rc = pbpb_config()
    nx, ny, nz = int(ic[0]), int(ic[1]), int(ic[2])
    pa, pb, pc = glauber_event(seed=5)
    g = EchoGrid(ic, rc, eos_rows=read_eos_rows())
    assert g.init_glauber(GLAUBER_PAR, pa, pb, pc) == 0
    print("\nfull size 384x384x256, unfloored Glauber state: step 1", flush=True)
    g.steps(1)
    assert g.scalars()["crashed"] == 0 and g.repairs() == 0
    dt1 = g.scalars()["dt"]
    v1, _ = g.state()
    print("step 2", flush=True)
    with pytest.raises(EchoError) as ei:
        g.steps(1)
    msg = str(ei.value)
    print(msg, f"repaired cells: {g.repairs()}", flush=True)
    assert ei.value.code == 1 and "repair failed" in msg
    m = re.search(r"ix=(\d+) iy=(\d+) iz=(\d+)", msg)
    cell = (int(m.group(1)) - 1, int(m.group(2)) - 1, int(m.group(3)) - 1)
    dt2 = g.scalars()["dt"]
    g.close()
    # the oracle's box around the failing cell
    H = 26
    lo = [min(max(c - H, 0), n - 2 * H) for c, n in zip(cell, (nx, ny, nz))]
    x1, y1, z1 = rc[0], rc[2], rc[4]
    dx, dy, dz = (rc[1] - rc[0]) / nx, (rc[3] - rc[2]) / ny, (rc[5] - rc[4]) / nz
    icb, rcb = make_config(nx=2 * H, ny=2 * H, nz=2 * H, coordinates=2, viscosity=1, bulk=0, eqstate=3,
                           rec="WENO5", nrk=3, bc=2, obtained_zz=1, init_type=3,
                           xlim=(x1 + lo[0] * dx, x1 + (lo[0] + 2 * H) * dx),
                           ylim=(y1 + lo[1] * dy, y1 + (lo[1] + 2 * H) * dy),
                           zlim=(z1 + lo[2] * dz, z1 + (lo[2] + 2 * H) * dz),
                           tmin=rc[6], cfl=rc[7], maxdt=dt1, eta_over_s=rc[10], tau_pi_coeff=rc[11],
                           enezero=rc[12])
    o, v0 = _glauber_state(icb, rcb, pa, pb, pc)
    assert o.init_from_prim(v0) == 0
    assert o.steps(1) == 0 and o.scalars()["dt"] == dt1
    vo, _ = o.state()
    inner = []
    for d, n in zip(range(3), (nx, ny, nz)):
        a = 0 if lo[d] == 0 else 9
        b = 2 * H if lo[d] + 2 * H == n else 2 * H - 9
        inner.append(slice(a, b))
    box = v1[:, lo[2]:lo[2] + 2 * H, lo[1]:lo[1] + 2 * H, lo[0]:lo[0] + 2 * H]
    sel = (slice(None), inner[2], inner[1], inner[0])
    err = rel_linf(box[sel], vo[sel])
    print(f"box at {lo}: step 1 rel L-inf {err:.2e}", flush=True)
    assert err <= 1e-10
    o.set_maxdt(dt2)
    assert o.steps(1) == 1 and o.repair_counts()["u2v_fail"] == 1
    c = o.crash_cell()
    got = (lo[0] + c % (2 * H), lo[1] + (c // (2 * H)) % (2 * H), lo[2] + c // (4 * H * H))
    assert got == cell, (got, cell)


# ---------------------------------------------------------------------------
# The viscous resistive system (nv = 19, IMEX-SSP3(4,3,3)) at the two large
# single-event sizes.  The reflection above is the rotation by pi about the
# y axis, under which B, E transform like v and psi (div B) is even.
# ---------------------------------------------------------------------------
PARITY19 = np.concatenate([PARITY, [-1, 1, -1, -1, 1, -1, 1]])
GROUPS19 = GROUPS + [[12, 13, 14], [15, 16, 17], [18]]


def _bench():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
    import bench
    return bench


def _symmetric_resistive_run(nx, ny, nz, dx, deta, nucleus, b, eta_over_s, steps, bulk_peak=None):
    from libs import make_config
    from echoamd import EchoGrid, read_eos_rows
    bench = _bench()
    ic, rc = make_config(nx=nx, ny=ny, nz=nz, coordinates=2, viscosity=1, bulk=1 if bulk_peak else 0,
                         eqstate=3, mhd=2,
                         divclean=1, algo_solver=7, rec="WENO5", nrk=4, bc=2, obtained_zz=1, init_type=3,
                         xlim=(-nx * dx / 2, nx * dx / 2), ylim=(-ny * dx / 2, ny * dx / 2),
                         zlim=(-nz * deta / 2, nz * deta / 2), tmin=0.4, cfl=0.4, maxdt=0.1,
                         eta_over_s=eta_over_s, tau_pi_coeff=5.0, enezero=0.001, sigma_el=0.0058,
                         sigma_chi=0.0)
    eos = read_eos_rows()
    v0, _ = bench.workload_state(ic, rc, eos, seed=11, b=b, nucleus=nucleus)
    v0 = 0.5 * (v0 + v0[:, ::-1, :, ::-1] * PARITY19[:, None, None, None])
    assert np.abs(v0[13]).max() > 0.05   # the dipole field is there
    print(f"\nresistive {nx}x{ny}x{nz}: symmetric initial state and {steps} IMEX steps", flush=True)
    g = EchoGrid(ic, rc, eos_rows=eos)
    assert g.nv == 19
    if bulk_peak:
        g.set_bulk_peak(*bulk_peak)
    assert g.init_from_prim(v0) == 0
    del v0
    g.steps(steps)
    s = g.scalars()
    assert s["crashed"] == 0 and s["iter"] == steps
    v, _ = g.state()
    g.close()
    assert np.all(np.isfinite(v))
    mirrored = v[:, ::-1, :, ::-1] * PARITY19[:, None, None, None]
    worst = {}
    for grp in GROUPS19:
        scale = max(np.abs(v[k]).max() for k in grp)
        if scale == 0.0:
            assert all(np.array_equal(v[k], mirrored[k]) for k in grp)
            continue
        worst[grp[0]] = max(float(np.abs(v[k] - mirrored[k]).max() / scale) for k in grp)
    # flow, currents and cleaning are all non-trivial
    assert np.abs(v[1]).max() > 1e-3 and np.abs(v[15:18]).max() > 1e-6 and np.abs(v[18]).max() > 0
    print("max reflection asymmetry / scale by group (first index): "
          + " ".join(f"{k}:{w:.1e}" for k, w in worst.items()), flush=True)
    return worst


def _check_symmetry(worst):
    """1e-10 of the scale for the fluid, B and E (measured: at most 2.3e-11);
    1e-8 for the cleaning scalar psi, whose scale is itself the size of the
    discrete div B of the smooth dipole, so that the build's rounding
    asymmetry shows larger against it (measured 7.6e-11 and 5.0e-10).  A
    wrong index, sign or stencil offset gives O(1)."""
    assert max(w for k, w in worst.items() if k != 18) <= 1e-10, worst
    assert worst.get(18, 0.0) <= 1e-8, worst


def test_resistive_reflection_symmetry_192x192x128():
    """The single-event Au-Au grid of the benchmark (dx 0.15, deta 0.1,
    eta/s 0.1, sigma_el 5.8 MeV, dipole B), 5 IMEX steps."""
    _check_symmetry(_symmetric_resistive_run(192, 192, 128, 0.15, 0.1, _bench().AU197, 7.0, 0.1, 5))


def test_resistive_reflection_symmetry_384x384x256():
    """The large Pb-Pb grid (dx 0.08, deta 0.05, b = 5 fm, eta/s 0.12,
    zeta/s peaked at T_c = 155 MeV with width 20 MeV and height 0.08),
    3 IMEX steps."""
    worst = _symmetric_resistive_run(384, 384, 256, 0.08, 0.05, _bench().PB208, 5.0, 0.12, 3,
                                     bulk_peak=(0.08, 0.155, 0.02))
    assert 5 in worst   # the bulk pressure developed
    _check_symmetry(worst)
