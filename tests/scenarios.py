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


# ---------------------------------------------------------------------------
# States that drive the repair branches of evolve_df{x,y,z} (evolve.f90:
# 334-369) and evolve_u2v (evolve.f90:795-842)
# ---------------------------------------------------------------------------
def rough_patch(nx=12, ny=10, nz=8, eqstate=1, rec="WENO5", seed=1, amp=0.9, pr_lo=-6.0):
    """fireball3d with a seeded rough patch inside: velocities in random
    directions up to |v| = amp and pressures spread over 10**pr_lo .. 1 times
    the smooth value.  Its reconstructions overshoot (p <= 0, v^2 >= 1) and
    its cons2prim fails in places, so both repairs run."""
    rng = np.random.default_rng(seed)
    ic, rc, v = fireball3d(nx, ny, nz, eqstate=eqstate, rec=rec, seed=seed)
    sl = (slice(2, nz - 2), slice(3, ny - 3), slice(3, nx - 3))
    shp = v[0][sl].shape
    d = rng.standard_normal((3,) + shp)
    d /= np.linalg.norm(d, axis=0)
    sp = amp * rng.uniform(0.0, 1.0, shp) ** 0.3
    v[1][sl] = d[0] * sp
    v[2][sl] = d[1] * sp
    v[3][sl] = d[2] * sp
    v[4][sl] = v[4][sl] * 10 ** rng.uniform(pr_lo, 0, shp)
    return ic, rc, v


# interior cells (iz, iy, ix) whose conserved state is made superluminal; the
# first two are x neighbours, the third is the first's y neighbour, so the
# later cells' repairs read the earlier cells' repaired values
FORCED_CELLS = ((4, 5, 5), (4, 5, 6), (4, 6, 5))


def forced_background(eqstate=2, viscosity=0, rec="TVD2", pbg=1e-3, coordinates=1, seed=0,
                      smooth_temp=0.0, nx=10, ny=10, nz=8):
    """A smooth, cold background (p/rho ~ pbg) for host-forced failures."""
    rng = np.random.default_rng(seed)
    ic, rc = make_config(nx=nx, ny=ny, nz=nz, coordinates=coordinates, viscosity=viscosity, bulk=1,
                         eqstate=eqstate, rec=rec, nrk=3, bc=2, xlim=(-1, 1), ylim=(-1, 1),
                         zlim=(-1, 1), tmin=1.0, maxdt=1.0, init_type=0, smooth_temp=smooth_temp)
    nv = 12 if viscosity else 5
    x = cell_centres(nx, -1, 1)
    y = cell_centres(ny, -1, 1)
    z = cell_centres(nz, -1, 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    v = np.zeros((nv, nz, ny, nx))
    v[0] = 1.0 + 0.1 * np.cos(X)
    v[1] = 0.1 * np.sin(2 * Y)
    v[2] = 0.1 * np.sin(2 * X)
    v[3] = 0.05 * np.sin(Z)
    v[4] = pbg * (1.0 + 0.3 * np.sin(3 * X + Y)) * (1 + 0.05 * rng.standard_normal(X.shape))
    if viscosity:
        v[9] = 0.1 * v[4]
        v[10] = 0.1 * v[4]
        v[6] = 0.02 * v[4]
    return ic, rc, v


def force_superluminal(u, cells=FORCED_CELLS, factor=3.0):
    """u with S_x = factor * E at the given cells (|S| > E: no primitive
    state has these conserved variables, so cons2prim fails there)."""
    u = u.copy()
    for (z, y, x) in cells:
        u[1, z, y, x] = factor * u[4, z, y, x]
    return u


def fast_eta(rec="TVD2", v_eta=0.99):
    """Bjorken grid with two cells moving at v^eta = v_eta at tau = 1: as tau
    grows, their own v^2 = tau^2 (v^eta)^2 passes 1, which drives the
    maxspeed rescaling of the reconstruction repair (evolve.f90:342-346) and,
    one step later, a failed con2prim repair."""
    ic, rc, v = forced_background(eqstate=1, rec=rec, pbg=0.05, coordinates=2)
    for (z, y, x) in ((4, 5, 5), (3, 4, 6)):
        v[1, z, y, x] = 0.05
        v[2, z, y, x] = 0.0
        v[3, z, y, x] = v_eta
    return ic, rc, v


# ---------------------------------------------------------------------------
# Glauber-MC events (test-side sampler, numpy's seeded generator)
# ---------------------------------------------------------------------------
# Glauber-MC deposition parameters (echo_glauber_params)
GLAUBER_PAR = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=5.36, deta=3.0, sigeta=0.8, rhocenter=1.0)


def woods_saxon_nucleus(rng, A, R, a):
    """(x, y) of A nucleons: radii from r^2 / (1 + exp((r - R) / a)) by
    rejection, isotropic directions, centre of mass at the origin
    (the sampling of glaubermc.f90:251-307 with another generator)."""
    rmax = R + 10 * a
    pts = []
    while len(pts) < A:
        r = rng.uniform(0, rmax, 8 * A)
        keep = rng.uniform(0, 1, r.size) * rmax ** 2 < r ** 2 / (1 + np.exp((r - R) / a))
        pts.extend(r[keep].tolist())
    r = np.array(pts[:A])
    cth = rng.uniform(-1, 1, A)
    phi = rng.uniform(0, 2 * np.pi, A)
    sth = np.sqrt(1 - cth ** 2)
    p = np.stack([r * sth * np.cos(phi), r * sth * np.sin(phi)], 1)
    return p - p.mean(0)


def glauber_event(seed=5, A=208, R=6.62, a=0.546, sigma_nn=6.76, b=5.0):
    """Participants of nuclei A (at +b/2) and B (at -b/2) and the binary
    collision points (midpoints of colliding pairs, d^2 <= sigma_nn / pi as
    glaubermc.f90:408-420); defaults: Pb-Pb, sigma_NN = 67.6 mb, b = 5 fm."""
    rng = np.random.default_rng(seed)
    na = woods_saxon_nucleus(rng, A, R, a) + [b / 2, 0]
    nb = woods_saxon_nucleus(rng, A, R, a) - [b / 2, 0]
    hit = ((na[:, None] - nb[None]) ** 2).sum(-1) <= sigma_nn / np.pi
    ia, ib = np.nonzero(hit)
    return na[hit.any(1)], nb[hit.any(0)], 0.5 * (na[ia] + nb[ib])


def pbpb_config(nx=384, ny=384, nz=256, zc=0.0, dx=0.08, deta=0.05, eta_over_s=0.12):
    """The large single-event grid: 384 x 384 x 256, dx = 0.08 fm,
    deta = 0.05, tau0 = 0.4 fm, viscous hydro, tabulated EoS, WENO5 + RK3,
    ENEZERO 0.001; nz and the centre zc select an eta slab of it."""
    return make_config(nx=nx, ny=ny, nz=nz, coordinates=2, viscosity=1, bulk=0, eqstate=3,
                       rec="WENO5", nrk=3, bc=2, obtained_zz=1, init_type=3,
                       xlim=(-nx * dx / 2, nx * dx / 2), ylim=(-ny * dx / 2, ny * dx / 2),
                       zlim=(zc - nz * deta / 2, zc + nz * deta / 2), tmin=0.4, cfl=0.4, maxdt=0.1,
                       eta_over_s=eta_over_s, tau_pi_coeff=5.0, enezero=0.001)


# ---------------------------------------------------------------------------
# Resistive MHD (mhd = 2, with the Dedner scalar), with and without viscosity
# ---------------------------------------------------------------------------
def resistive_layout(viscosity):
    """(index of B^x, index of E^x, index of psi, nv)"""
    nf = 12 if viscosity else 5
    return nf, nf + 3, nf + 6, nf + 7


def rmhd_fireball(nx=16, ny=14, nz=10, viscosity=1, eqstate=1, nrk=4, B0=0.3, E0=0.05, coordinates=2,
                  sigma_el=0.0058, sigma_chi=0.0015, rec="WENO5", maxdt=1.0, bc=2, seed=0, smooth_temp=0.0,
                  obtained_zz=1):
    """fireball3d with a transverse magnetic field and an electric field
    (not the ideal -v x B, so that the current flows), for the resistive
    system; viscous runs carry the fireball's shear tensor."""
    ic, rc, vf = fireball3d(nx, ny, nz, viscosity=viscosity, eqstate=eqstate, rec=rec, nrk=nrk,
                            coordinates=coordinates, bc=bc, seed=seed, smooth_temp=smooth_temp,
                            obtained_zz=obtained_zz)
    ic[6] = 2
    ic[7] = 1
    ic[12] = 7
    rc[8] = maxdt
    rc[14] = sigma_el
    rc[15] = sigma_chi
    kb, ke, kg, nv = resistive_layout(viscosity)
    v = np.zeros((nv,) + vf.shape[1:])
    v[:kb] = vf
    x = cell_centres(nx, -3, 3)
    y = cell_centres(ny, -3, 3)
    z = cell_centres(nz, -1, 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    v[kb + 1] = B0 * np.exp(-(X ** 2 + Y ** 2) / 8)
    v[kb] = 0.2 * B0 * np.sin(Y)
    v[kb + 2] = 0.05 * B0
    v[ke] = E0 * np.exp(-(X ** 2 + Y ** 2) / 6) * np.cos(0.5 * Y)
    v[ke + 1] = 0.5 * E0 * np.sin(0.5 * X)
    v[ke + 2] = 0.1 * E0 * np.exp(-(X ** 2 + Y ** 2) / 8)
    return ic, rc, v


def reference_resistive_config(ic):
    """The reference's set-up of a resistive grid for its writers: it has no
    inversion for the resistive system (no algorithm 7, system.f90:1160-1162;
    its set-up stops on ALG_SOLV 7), so it is set up with algorithm 5, which
    out_vars, out_dump and the hypersurface finder never call, and is then
    given the oracle's states."""
    ic = ic.copy()
    ic[12] = 5
    return ic


def uniform_resistive(E0=(0.3, -0.2, 0.1), B0=(0.0, 0.0, 0.0), p=1.0, sigma_el=0.0058, nrk=4, n=4,
                      maxdt=0.01, coordinates=1, viscosity=0):
    """A uniform fluid at rest with uniform E and B on a small periodic grid:
    no gradients, so only the local current acts (Minkowski: dE/dt =
    -sigma (E + v x B ...) with v = 0, i.e. E(t) = E0 exp(-sigma t) for
    B = 0, the energy moving to the fluid)."""
    ic, rc = make_config(nx=n, ny=n, nz=n, coordinates=coordinates, viscosity=viscosity, mhd=2, divclean=1,
                         eqstate=1, algo_solver=7, nrk=nrk, bc=0, init_type=0,
                         xlim=(-1.0, 1.0), ylim=(-1.0, 1.0), zlim=(-1.0, 1.0), tmin=1.0,
                         cfl=0.4, maxdt=maxdt, sigma_el=sigma_el, sigma_chi=0.0)
    kb, ke, kg, nv = resistive_layout(viscosity)
    v = np.zeros((nv, n, n, n))
    v[0] = 1.0
    v[4] = p
    for k in range(3):
        v[ke + k] = E0[k]
        v[kb + k] = B0[k]
    return ic, rc, v


def light_wave(n=64, amp=1e-3, nrk=3, rec="WENO5", sigma_el=0.0):
    """A small electromagnetic plane wave (E^y = B^eta = amp sin(2 pi x)) in a
    dense uniform fluid at rest, Minkowski, periodic in x: with sigma = 0 it
    travels in +x at the speed of light."""
    ic, rc = make_config(nx=n, ny=1, nz=1, coordinates=1, mhd=2, divclean=1, eqstate=1, algo_solver=7,
                         rec=rec, nrk=nrk, bc=0, init_type=0, xlim=(-0.5, 0.5), ylim=(-0.5, 0.5),
                         zlim=(-0.5, 0.5), tmin=0.0, cfl=0.4, maxdt=1.0, sigma_el=sigma_el, sigma_chi=0.0)
    kb, ke, kg, nv = resistive_layout(0)
    x = cell_centres(n, -0.5, 0.5)
    v = np.zeros((nv, 1, 1, n))
    v[0] = 1.0
    v[4] = 100.0
    v[ke + 1, 0, 0] = amp * np.sin(2 * np.pi * x)
    v[kb + 2, 0, 0] = amp * np.sin(2 * np.pi * x)
    return ic, rc, v


def forced_resistive(viscosity=0, rec="TVD2", pbg=1e-2, nrk=4, sigma_el=0.0058, B0=0.05, E0=0.02):
    """forced_background as a resistive grid (small B and E fields); the
    caller forces superluminal conserved states with force_superluminal, so
    that the implicit solves of the IMEX stages fail and are repaired."""
    ic, rc, vf = forced_background(eqstate=1, viscosity=viscosity, rec=rec, pbg=pbg)
    ic = ic.copy()
    rc = rc.copy()
    ic[6], ic[7], ic[10], ic[12] = 2, 1, nrk, 7
    rc[14] = sigma_el
    kb, ke, kg, nv = resistive_layout(viscosity)
    v = np.zeros((nv,) + vf.shape[1:])
    v[:kb] = vf
    x = cell_centres(vf.shape[3], -1, 1)
    y = cell_centres(vf.shape[2], -1, 1)
    z = cell_centres(vf.shape[1], -1, 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    v[kb] = B0 * np.cos(Y)
    v[kb + 1] = B0 * np.sin(X)
    v[ke + 2] = E0 * np.cos(X + Y)
    return ic, rc, v
