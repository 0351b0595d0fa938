"""The grid update at the largest configured size, 384 x 384 x 256 cells
(viscous hydro, tabulated EoS, WENO5 + HLL + RK3, Milne coordinates), where
the oracle is too slow, through a size-independent property: a state
symmetric under the reflection (x, eta) -> (-x, -eta) stays symmetric.

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
