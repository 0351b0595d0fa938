"""The oracle's initial magnetic field (oracle_init.c, orc_glauber_bfield,
restating generate_B_field_profile, glaubermc.f90:1009-1168), pinned where
physics fixes it:
- without conductivity the classical field of one proton is the boosted
  Coulomb field of a point charge, B = gamma u x E' (E' the rest-frame
  field), with u along +z for nucleus A and -z for nucleus B;
- with electric conductivity the classical field solves Maxwell's equations
  in a conducting medium away from the charge, curl curl B + sigma dB/dt +
  d2B/dt2 = 0 with sigma = sigma_el / hbar (finite differences); the same
  field without the d2B/dt2 term leaves an O(1) residual;
- bookkeeping: participants weighted by pw, lists read up to their first
  neutron, superposition, magfield_type 2 (classical only) azimuthal and 3
  (chiral only) without azimuthal part.
The chiral fields Br, Bz are restated, not pinned: the reference's formulas
(glaubermc.f90:1047-1050) are not divergence-free (a finite-difference
divergence of the order of the field's gradients), so no property of a
physical field constrains them; parity unpinned for those terms.
"""
import numpy as np

from libs import cell_centres, make_config, oracle_glauber_bfield

HBAR, QBIG, SQRH = 0.197326, 0.30282212, 0.444214   # common.f90:51-55
BP = dict(gamma_col=20.0, b_amplify=1.0, pw=0.5, magfield_type=1)


def _grid(oracle, xlim, ylim, zlim, n, t, sigma_el, sigma_chi, coordinates=1):
    ic, rc = make_config(nx=n[0], ny=n[1], nz=n[2], coordinates=coordinates, mhd=1, divclean=0,
                         algo_solver=5, eqstate=1, xlim=xlim, ylim=ylim, zlim=zlim, tmin=t,
                         sigma_el=sigma_el, sigma_chi=sigma_chi)
    oracle.setup(ic, rc)
    return ic, rc


def _field(oracle, nucleons, bp=BP, xlim=(-2.0, 2.5), ylim=(-1.5, 2.0), zlim=(-0.8, 0.6), n=(6, 5, 4),
           t=0.4, sigma_el=0.0, sigma_chi=0.0, coordinates=1):
    _grid(oracle, xlim, ylim, zlim, n, t, sigma_el, sigma_chi, coordinates)
    v = np.zeros((oracle.nv, n[2], n[1], n[0]))
    assert oracle_glauber_bfield(oracle, bp, nucleons, v) == 0
    return v[5:8]   # B^x, B^y, B^eta (ideal MHD without GLM)


def _lists(a_spec=(), a_part=(), b_spec=(), b_part=()):
    out = []
    for L in (a_spec, a_part, b_spec, b_part):
        L = list(L)
        xy = np.array([p[:2] for p in L], dtype=np.float64).reshape(-1, 2)
        ty = np.array([p[2] if len(p) > 2 else 0 for p in L], dtype=np.int32)
        out.append((xy, ty))
    return out


def _coords(xlim=(-2.0, 2.5), ylim=(-1.5, 2.0), zlim=(-0.8, 0.6), n=(6, 5, 4)):
    x, y, z = cell_centres(n[0], *xlim), cell_centres(n[1], *ylim), cell_centres(n[2], *zlim)
    return np.meshgrid(z, y, x, indexing="ij")


def test_coulomb_field_of_a_moving_proton(oracle):
    gam, t, xs, ys = BP["gamma_col"], 0.4, 0.3, -0.2
    vel = np.sqrt(1 - 1 / gam ** 2)
    Z, Y, X = _coords()
    for nucleus in ("A", "B"):
        L = _lists(a_spec=[(xs, ys)]) if nucleus == "A" else _lists(b_spec=[(xs, ys)])
        B = _field(oracle, L, t=t)
        # rest frame of the charge: longitudinal distance gamma * zpos, with
        # zpos = t v + z (A) or t v - z (B) as the reference measures it
        zpos = t * vel + Z if nucleus == "A" else t * vel - Z
        Ep = np.stack([X - xs, Y - ys, gam * zpos], axis=-1)
        Ep = QBIG / (4 * np.pi) * Ep / np.linalg.norm(Ep, axis=-1)[..., None] ** 3
        u = np.array([0.0, 0.0, vel if nucleus == "A" else -vel])
        Bl = gam * np.cross(u, Ep) * SQRH * BP["b_amplify"]   # sqrt(hbar c): the reference's units
        for k in range(3):
            assert np.allclose(B[k], Bl[..., k], rtol=1e-13, atol=1e-16 * np.abs(Bl).max()), (nucleus, k)


def _maxwell_residual(oracle, x0, y0, z0, t0, sigma_el, h=2e-3, drop_tt=False):
    L = _lists(a_spec=[(0.1, -0.2)])
    bp = dict(BP, magfield_type=2)
    lim = lambda c: (c - 1.5 * h, c + 1.5 * h)
    args = dict(xlim=lim(x0), ylim=lim(y0), zlim=lim(z0), n=(3, 3, 3), sigma_el=sigma_el)
    B0 = _field(oracle, L, bp, t=t0, **args)
    Bp = _field(oracle, L, bp, t=t0 + h, **args)[:, 1, 1, 1]
    Bm = _field(oracle, L, bp, t=t0 - h, **args)[:, 1, 1, 1]
    c = B0[:, 1, 1, 1]
    lap = (B0[:, 1, 1, 2] + B0[:, 1, 1, 0] + B0[:, 1, 2, 1] + B0[:, 1, 0, 1] + B0[:, 2, 1, 1] +
           B0[:, 0, 1, 1] - 6 * c) / h ** 2
    dt = (Bp - Bm) / (2 * h)
    dtt = (Bp + Bm - 2 * c) / h ** 2
    s = sigma_el / HBAR
    # the classical field is azimuthal, so div B = 0 and curl curl B = -lap B
    r = -lap + s * dt + (0.0 if drop_tt else dtt)
    scale = np.abs(lap) + np.abs(s * dt) + np.abs(dtt)
    return np.abs(r).max() / scale.max()


def test_classical_field_solves_maxwell_in_a_conductor(oracle):
    for sigma_el in (0.0, 0.0058, 0.05):
        for pt in ((1.0, 0.5, 0.3, 0.2), (2.0, -1.0, -0.2, 0.4), (-0.7, 1.3, 0.1, 0.6)):
            r = _maxwell_residual(oracle, *pt, sigma_el)
            assert r < 1e-4, (sigma_el, pt, r)
            # the diffusion form (no second time derivative) does not hold
            assert _maxwell_residual(oracle, *pt, sigma_el, drop_tt=True) > 0.1


def test_bookkeeping(oracle):
    p1, p2 = (0.4, -0.3), (-1.1, 0.8)
    kw = dict(sigma_el=0.0058, sigma_chi=0.0015, t=0.5)
    # each list is read up to its first neutron
    B1 = _field(oracle, _lists(a_spec=[p1]), **kw)
    Bc = _field(oracle, _lists(a_spec=[p1, (0.0, 0.0, 1), p2]), **kw)
    assert np.array_equal(B1, Bc)
    # superposition
    B2 = _field(oracle, _lists(a_spec=[p2]), **kw)
    B12 = _field(oracle, _lists(a_spec=[p1, p2]), **kw)
    assert np.allclose(B12, B1 + B2, rtol=1e-13, atol=1e-15 * np.abs(B12).max())
    # participants weigh pw
    Bp = _field(oracle, _lists(a_part=[p1]), **kw)
    assert np.allclose(Bp, BP["pw"] * B1, rtol=1e-13, atol=1e-16)
    # classical only: azimuthal around the charge, no longitudinal part;
    # chiral only: no azimuthal part
    Z, Y, X = _coords()
    dx, dy = X - p1[0], Y - p1[1]
    for typ in (2, 3):
        B = _field(oracle, _lists(b_spec=[p1]), dict(BP, magfield_type=typ), **kw)
        radial = B[0] * dx + B[1] * dy
        azim = -B[0] * dy + B[1] * dx
        if typ == 2:
            assert np.abs(radial).max() <= 1e-12 * np.abs(azim).max()
            assert np.abs(B[2]).max() == 0.0
        else:
            assert np.abs(azim).max() <= 1e-12 * np.abs(radial).max()
            assert np.abs(B[2]).max() > 0.0


def test_bjorken_plane_eta_zero_is_the_lab_field(oracle):
    """At eta = 0 the Milne cell sits at z = 0, t = tau0 with unit
    conversion factors for B^x, B^y (glaubermc.f90:1017-1022)."""
    L = _lists(a_spec=[(0.4, -0.3)], b_part=[(-0.6, 0.2)])
    kw = dict(zlim=(-0.05, 0.05), n=(6, 5, 1), t=0.6, sigma_el=0.0058, sigma_chi=0.0)
    Bm = _field(oracle, L, coordinates=2, **kw)
    Bl = _field(oracle, L, coordinates=1, **kw)
    assert np.allclose(Bm[:2], Bl[:2], rtol=1e-14, atol=0)


def test_resistive_glauber_state_has_the_ideal_fields():
    """The Glauber state of a resistive grid (init.f90:2732-2832 fills
    kbx..kbz; the layout of system.f90:92-124 puts B, E, psi after the fluid
    variables): the fluid variables and B equal those of the ideal-MHD (or,
    with viscosity, the viscous hydro) grid of the same event, E and psi are
    zero, and the state initialises (prim2cons, init.f90:2937-2958)."""
    from libs import Oracle, oracle_glauber_prim
    from scenarios import resistive_layout
    rng = np.random.default_rng(3)
    pa = rng.normal(0, 2.0, (40, 2)) + [1.0, 0.0]
    pb = rng.normal(0, 2.0, (36, 2)) - [1.0, 0.0]
    pc = rng.normal(0, 1.6, (90, 2))
    lists = [(rng.normal(0, 2.2, (20, 2)) + [s, 0.0], np.r_[np.zeros(9, np.int32), np.ones(11, np.int32)])
             for s in (4.0, 1.0, -4.0, -1.0)]
    par = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=5.36, deta=3.0, sigeta=0.8, rhocenter=1.0)
    bp = dict(gamma_col=106.6, b_amplify=0.1, pw=0.5, magfield_type=1)
    box = dict(nx=16, ny=14, nz=6, coordinates=2, eqstate=1, xlim=(-6.0, 6.0), ylim=(-5.25, 5.25),
               zlim=(-3.0, 3.0), tmin=0.6, init_type=3, eta_over_s=0.1, sigma_el=0.0058, sigma_chi=0.0015)
    o = Oracle()

    def state(**kw):
        ic, rc = make_config(**box, **kw)
        o.setup(ic, rc)
        e, v = oracle_glauber_prim(o, par, pa, pb, pc)
        assert e == 0
        if kw.get("mhd"):
            assert oracle_glauber_bfield(o, bp, lists, v) == 0
        return v, o.init_from_prim(v)

    v_ideal, e = state(mhd=1, divclean=1, algo_solver=5)
    assert e == 0
    v_visc, e = state(viscosity=1)
    assert e == 0
    for viscosity in (0, 1):
        v, e = state(mhd=2, divclean=1, algo_solver=7, nrk=4, viscosity=viscosity)
        assert e == 0
        kb, ke, kg, nv = resistive_layout(viscosity)
        assert v.shape[0] == nv
        assert np.array_equal(v[:5], v_ideal[:5])
        assert np.array_equal(v[kb:kb + 3], v_ideal[5:8])
        assert np.abs(v[kb:kb + 3]).max() > 0.0
        assert not v[ke:].any()
        if viscosity:
            assert np.array_equal(v[:12], v_visc)
