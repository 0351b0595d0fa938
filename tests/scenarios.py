"""Seeded, deterministic test scenarios shared by the oracle and GPU tests.

Each returns (ic, rc, v0): configuration arrays (oracle/ref_harness.f90
layout) and the initial primitives with shape (nv, nz, ny, nx).
"""
import numpy as np

from libs import HBAR, cell_centres, gubser_prim, make_config

# Gubser flow at T0 = 1.2 GeV at r = 0, tau0 = q = 1 (in fm units T^ = T0/hbar)
GUBSER_THAT0 = 1.2 / HBAR


def gubser2d(n=128, rec="WENO5", nrk=3, dx=0.1, cfl=0.4, that0=GUBSER_THAT0):
    """Ideal Gubser flow, Bjorken coordinates, p = e/3, n x n x 1 cells of
    size dx, outflow boundaries (init.f90:2901-2961 with T0 -> that0)."""
    half = n * dx / 2
    ic, rc = make_config(nx=n, ny=n, nz=1, coordinates=2, rec=rec, nrk=nrk, bc=2,
                         eqstate=1, init_type=4, xlim=(-half, half), ylim=(-half, half),
                         zlim=(-0.5, 0.5), tmin=1.0, cfl=cfl, maxdt=1.0)
    v0 = gubser_prim(n, n, (-half, half), (-half, half), 1.0, that0=that0)
    return ic, rc, v0


def fireball3d(nx=24, ny=20, nz=16, viscosity=0, eqstate=1, rec="WENO5", nrk=3,
               coordinates=2, bc=2, smooth_temp=0.0, bulk=1, obtained_zz=1, seed=0,
               tmin=1.0):
    """A smooth 3-D fireball with transverse and longitudinal flow and small
    seeded fluctuations; viscous runs start with a non-zero shear tensor."""
    rng = np.random.default_rng(seed)
    ic, rc = make_config(nx=nx, ny=ny, nz=nz, coordinates=coordinates, viscosity=viscosity,
                         bulk=bulk, eqstate=eqstate, rec=rec, nrk=nrk, bc=bc,
                         xlim=(-3.0, 3.0), ylim=(-3.0, 3.0), zlim=(-1.0, 1.0),
                         init_type=0, smooth_temp=smooth_temp, obtained_zz=obtained_zz,
                         tmin=tmin, maxdt=1.0)
    nv = 12 if viscosity else 5
    x = cell_centres(nx, -3, 3)
    y = cell_centres(ny, -3, 3)
    z = cell_centres(nz, -1, 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tab = eqstate == 3
    e = 5.0 * np.exp(-(X ** 2 + Y ** 2) / 4 - Z ** 2) + (0.5 if tab else 0.01)
    e = e * (1.0 + 0.02 * rng.standard_normal(e.shape))
    v = np.zeros((nv, nz, ny, nx))
    v[0] = 1.0
    v[1] = 0.1 * np.sin(X) * np.exp(-Y ** 2 / 8)
    v[2] = 0.1 * np.cos(Y) * np.exp(-X ** 2 / 8)
    v[3] = 0.05 * np.sin(Z) / tmin
    v[4] = e / 3.0 if not tab else e * 0.25
    if viscosity:
        v[9] = 0.01 * e     # pi^xx
        v[10] = 0.01 * e    # pi^yy
        v[6] = 0.002 * e * np.sin(X)   # pi^xy
    return ic, rc, v


def shock2d(n=48, rec="WENO5", nrk=3, eqstate=2):
    """Minkowski 2-D Riemann problem (four quadrants), ideal gas by default;
    exercises limiters and the reconstruction repairs."""
    ic, rc = make_config(nx=n, ny=n, nz=1, coordinates=1, rec=rec, nrk=nrk, bc=2,
                         eqstate=eqstate, init_type=1, xlim=(-0.5, 0.5), ylim=(-0.5, 0.5),
                         zlim=(-0.5, 0.5), tmin=0.0, cfl=0.4, maxdt=1.0)
    x = cell_centres(n, -0.5, 0.5)
    X, Y = np.meshgrid(x, x, indexing="xy")
    v = np.zeros((5, 1, n, n))
    hi = (X < 0) & (Y < 0)
    v[0, 0] = np.where(hi, 1.0, 0.1)
    v[4, 0] = np.where(hi, 1.0, 0.01)
    v[1, 0] = np.where((X >= 0) & (Y < 0), 0.5, 0.0)
    v[2, 0] = np.where((X < 0) & (Y >= 0), 0.5, 0.0)
    return ic, rc, v


def _groups(nv):
    """Variables measured on a common scale: the three components of the
    velocity (momentum) vector, and for MHD the field components together
    with the GLM scalar (same units)."""
    groups = [[0], [1, 2, 3], [4]]
    if nv in (8, 9):
        groups.append(list(range(5, nv)))
    else:
        groups += [[k] for k in range(5, nv)]
    return groups


def rel_linf(a, b):
    """max over variables of max|a-b| / scale, where the scale of a variable is
    max|b| over its group (see _groups); an all-zero group uses the absolute
    difference.  Vector components share the scale of their vector, so a
    component that is zero up to rounding (e.g. v^z of a 2-D run) is measured
    against the vector, not against its own rounding noise."""
    out = 0.0
    for grp in _groups(b.shape[0]):
        scale = max(np.abs(b[k]).max() for k in grp)
        for k in grp:
            d = np.abs(a[k] - b[k]).max()
            out = max(out, d / scale if scale > 0 else d)
    return out


def mhd2d(n=32, eqstate=1, algo=5, divclean=1, coordinates=1, rec="WENO5", nrk=3, bc=0, B0=0.5):
    """A smooth 2-D MHD field: a pressure bump advected by a rotating flow
    through an oblique magnetic field with a periodic perturbation."""
    ic, rc = make_config(nx=n, ny=n, nz=1, coordinates=coordinates, mhd=1, divclean=divclean,
                         eqstate=eqstate, algo_solver=algo, rec=rec, nrk=nrk, bc=bc, init_type=7,
                         xlim=(-1.0, 1.0), ylim=(-1.0, 1.0), zlim=(-0.5, 0.5), tmin=1.0,
                         cfl=0.4, maxdt=0.01, glm_alpha=5.0)
    nv = 8 + (1 if divclean else 0)
    x = cell_centres(n, -1, 1)
    X, Y = np.meshgrid(x, x, indexing="xy")
    r2 = X ** 2 + Y ** 2
    v = np.zeros((nv, 1, n, n))
    v[0, 0] = 1.0 + 0.2 * np.exp(-4 * r2)
    v[1, 0] = -0.3 * Y * np.exp(-2 * r2)
    v[2, 0] = 0.3 * X * np.exp(-2 * r2)
    v[4, 0] = 1.0 + 0.5 * np.exp(-8 * r2)
    v[5, 0] = B0 * (1 + 0.1 * np.sin(np.pi * Y))
    v[6, 0] = 0.5 * B0 * (1 + 0.1 * np.cos(np.pi * X))
    v[7, 0] = 0.1 * B0
    return ic, rc, v


def mhd_fireball3d(nx=16, ny=14, nz=10, algo=5, divclean=1, B0=0.3, eqstate=1):
    """3-D Bjorken fireball with a transverse magnetic field (p = e/3, or the
    tabulated EoS with algorithms 1/2)."""
    ic, rc, v5 = fireball3d(nx, ny, nz, eqstate=eqstate)
    ic[6] = 1
    ic[7] = divclean
    ic[12] = algo
    nv = 8 + (1 if divclean else 0)
    v = np.zeros((nv,) + v5.shape[1:])
    v[:5] = v5
    x = cell_centres(nx, -3, 3)
    y = cell_centres(ny, -3, 3)
    z = cell_centres(nz, -1, 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    v[6] = B0 * np.exp(-(X ** 2 + Y ** 2) / 8)
    v[5] = 0.2 * B0 * np.sin(Y)
    v[7] = 0.05 * B0
    return ic, rc, v


def mhd3d(nx=12, ny=10, nz=8, eqstate=3, algo=1, divclean=1, coordinates=2, B0=0.3):
    """A smooth 3-D MHD state (pressure bump, rotating flow, oblique field
    with a longitudinal modulation) on a Bjorken grid with outflow, for the
    tabulated EoS with algorithms 1/2."""
    ic, rc = make_config(nx=nx, ny=ny, nz=nz, coordinates=coordinates, mhd=1, divclean=divclean,
                         eqstate=eqstate, algo_solver=algo, bc=2, init_type=7,
                         xlim=(-1.0, 1.0), ylim=(-1.0, 1.0), zlim=(-0.5, 0.5), tmin=1.0,
                         cfl=0.4, maxdt=0.01, glm_alpha=5.0)
    nv = 8 + (1 if divclean else 0)
    x = cell_centres(nx, -1, 1)
    y = cell_centres(ny, -1, 1)
    z = cell_centres(nz, -0.5, 0.5)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r2 = X ** 2 + Y ** 2
    v = np.zeros((nv, nz, ny, nx))
    v[0] = 1.0 + 0.2 * np.exp(-4 * r2)
    v[1] = -0.3 * Y * np.exp(-2 * r2)
    v[2] = 0.3 * X * np.exp(-2 * r2)
    v[3] = 0.05 * np.sin(2 * np.pi * Z)
    v[4] = 1.0 + 0.5 * np.exp(-8 * r2 - 4 * Z ** 2)
    v[5] = B0 * (1 + 0.1 * np.sin(np.pi * Y))
    v[6] = 0.5 * B0 * (1 + 0.1 * np.cos(np.pi * X))
    v[7] = 0.1 * B0 * (1 + 0.2 * np.cos(2 * np.pi * Z))
    return ic, rc, v
