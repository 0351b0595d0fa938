"""The GPU Glauber-MC initial state (echo_init_glauber, and
echo_init_glauber_mhd with the protons' magnetic field) against the oracle
(oracle_init.c followed by the oracle's initialisation), and the first
steps evolved from it."""
import numpy as np
import pytest

from libs import make_config, oracle_glauber_bfield, oracle_glauber_prim
from scenarios import rel_linf

pytestmark = pytest.mark.gpu
PAR = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=5.36, deta=3.0, sigeta=0.8, rhocenter=1.0)


@pytest.mark.parametrize("viscosity,eqstate", [(1, 3), (0, 1), (0, 3)])
def test_glauber_initial_state(oracle, viscosity, eqstate):
    from echoamd import EchoGrid
    rng = np.random.default_rng(11)
    ic, rc = make_config(nx=40, ny=36, nz=16, coordinates=2, viscosity=viscosity, eqstate=eqstate,
                         xlim=(-6.0, 6.0), ylim=(-5.4, 5.4), zlim=(-4.0, 4.0), tmin=0.6,
                         init_type=3, eta_over_s=0.1, maxdt=0.05)
    pa = rng.normal(0, 2.0, (150, 2)) + [1.0, 0.0]
    pb = rng.normal(0, 2.0, (140, 2)) - [1.0, 0.0]
    pc = rng.normal(0, 1.6, (420, 2))
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, PAR, pa, pb, pc)
    assert e == 0
    assert oracle.init_from_prim(v0) == 0
    vo, uo = oracle.state()
    g = EchoGrid(ic, rc)
    assert g.init_glauber(PAR, pa, pb, pc) == 0
    vg, ug = g.state()
    # device exp and glibc exp differ in the last bit; the sums of 710
    # Gaussians keep that at the 1e-15 level
    assert rel_linf(vg, vo) <= 1e-12
    assert rel_linf(ug, uo) <= 1e-12
    assert oracle.steps(3) == 0
    g.steps(3)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= 1e-10 and rel_linf(ug, uo) <= 1e-10
    g.close()


def _nucleons(rng):
    """Spectator and participant lists of two nuclei, protons first as in the
    reference's files; the participants of B hold a neutron early, so the
    protons after it do not count."""
    out = []
    for n, npro, shift in ((60, 25, 4.0), (70, 30, 1.0), (60, 24, -4.0), (70, 28, -1.0)):
        xy = rng.normal(0, 2.2, (n, 2)) + [shift, 0.0]
        ty = np.r_[np.zeros(npro, np.int32), np.ones(n - npro, np.int32)]
        out.append((xy, ty))
    out[3][1][10] = 1
    return out


@pytest.mark.parametrize("coordinates,magfield_type,divclean", [(2, 1, 1), (1, 2, 0), (2, 3, 0)])
def test_glauber_mhd_initial_field(oracle, coordinates, magfield_type, divclean):
    """echo_init_glauber_mhd: energy deposition, the protons' field
    (generate_B_field_profile), then prim2cons, against the oracle; then
    three steps."""
    from echoamd import EchoGrid
    rng = np.random.default_rng(12)
    ic, rc = make_config(nx=32, ny=28, nz=12, coordinates=coordinates, mhd=1, divclean=divclean,
                         algo_solver=5, eqstate=1, xlim=(-6.0, 6.0), ylim=(-5.25, 5.25),
                         zlim=(-3.0, 3.0) if coordinates == 2 else (-1.5, 1.5), tmin=0.6, init_type=3,
                         maxdt=0.05, sigma_el=0.0058, sigma_chi=0.0015)
    pa = rng.normal(0, 2.0, (100, 2)) + [1.0, 0.0]
    pb = rng.normal(0, 2.0, (90, 2)) - [1.0, 0.0]
    pc = rng.normal(0, 1.6, (250, 2))
    nuc = _nucleons(rng)
    # the field scales with b_amplify; at 1.0 the near-vacuum corners of this
    # grid reach B^2/p ~ 30, where con2prim is so ill-conditioned that the
    # rounding differences of the two builds grow to 1e-3 within two steps;
    # at 0.1 they stay below B^2/p ~ 0.3
    bp = dict(gamma_col=106.6, b_amplify=0.1, pw=0.5, magfield_type=magfield_type)
    print(f"\nglauber mhd: coordinates={coordinates} magfield_type={magfield_type} divclean={divclean}",
          flush=True)
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, PAR, pa, pb, pc)
    assert e == 0
    assert oracle_glauber_bfield(oracle, bp, nuc, v0) == 0
    assert np.abs(v0[5:8]).max() > 0.0
    assert oracle.init_from_prim(v0) == 0
    vo, uo = oracle.state()
    g = EchoGrid(ic, rc)
    assert g.init_glauber_mhd(PAR, pa, pb, pc, bp, nuc) == 0
    vg, ug = g.state()
    # device and glibc atan2, sin, cos, exp and pow differ in the last bits;
    # the field is a sum over 107 protons
    assert rel_linf(vg, vo) <= 1e-12
    assert rel_linf(ug, uo) <= 1e-12
    assert oracle.steps(3) == 0
    g.steps(3)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= 1e-10 and rel_linf(ug, uo) <= 1e-10
    g.close()


def test_create_rejects_derivative_output():
    """The derivative output (out_derivatives) is not built: a configuration
    asking for it fails at creation instead of running without it."""
    from echoamd import EchoError, EchoGrid
    ic, rc = make_config(nx=8, ny=8, nz=4, coordinates=2, derivatives_out=1)
    with pytest.raises(EchoError):
        EchoGrid(ic, rc)


def test_glauber_rejects_bad_input():
    from echoamd import EchoError, EchoGrid
    ic, rc = make_config(nx=8, ny=8, nz=4, coordinates=2)
    g = EchoGrid(ic, rc)
    with pytest.raises(EchoError):
        g.init_glauber(dict(PAR, sig_mc=0.0), [[0, 0]], [[0, 0]], [[0, 0]])
    with pytest.raises(EchoError):
        g.init_glauber(PAR, np.zeros((3000, 2)), np.zeros((1000, 2)), np.zeros((200, 2)))
    nuc = [(np.zeros((1, 2)), np.zeros(1, np.int32))] * 4
    bp = dict(gamma_col=106.6, b_amplify=1.0, pw=0.5, magfield_type=1)
    with pytest.raises(EchoError):   # no magnetic field on this grid
        g.init_glauber_mhd(PAR, [[0, 0]], [[0, 0]], [[0, 0]], bp, nuc)
    g.close()
    ic, rc = make_config(nx=8, ny=8, nz=4, coordinates=2, mhd=1, algo_solver=5)
    g = EchoGrid(ic, rc)
    with pytest.raises(EchoError):
        g.init_glauber_mhd(PAR, [[0, 0]], [[0, 0]], [[0, 0]], dict(bp, magfield_type=4), nuc)
    with pytest.raises(EchoError):
        big = [(np.zeros((600, 2)), np.zeros(600, np.int32))] * 4
        g.init_glauber_mhd(PAR, [[0, 0]], [[0, 0]], [[0, 0]], bp, big)
    g.close()


def test_glauber_from_generated_event(oracle, tmp_path):
    """The whole initial-state path of the library: an Au-Au event of
    echo_glauber_generate (b = 6 fm), read back with echo_glauber_read_event,
    deposited on the device by echo_init_glauber, against the oracle's
    deposition of the same lists, and three steps."""
    from echoamd import EchoGrid, glauber_generate, glauber_read_event
    assert glauber_generate(tmp_path, 17, nucleus="Au", nconf=1, b=6.0) == 1
    ev = glauber_read_event(tmp_path, 1)
    assert len(ev["part_a"]) + len(ev["part_b"]) > 100 and len(ev["coll"]) > 100
    ic, rc = make_config(nx=48, ny=48, nz=12, coordinates=2, viscosity=1, eqstate=3,
                         xlim=(-9.6, 9.6), ylim=(-9.6, 9.6), zlim=(-3.0, 3.0), tmin=0.6,
                         init_type=3, eta_over_s=0.1, maxdt=0.05)
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, PAR, ev["part_a"], ev["part_b"], ev["coll"])
    assert e == 0
    assert oracle.init_from_prim(v0) == 0
    vo, uo = oracle.state()
    g = EchoGrid(ic, rc)
    assert g.init_glauber(PAR, ev["part_a"], ev["part_b"], ev["coll"]) == 0
    vg, ug = g.state()
    assert rel_linf(vg, vo) <= 1e-12 and rel_linf(ug, uo) <= 1e-12
    assert oracle.steps(3) == 0
    g.steps(3)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= 1e-10 and rel_linf(ug, uo) <= 1e-10
    g.close()


@pytest.mark.parametrize("viscosity", [0, 1])
def test_glauber_resistive_initial_state(oracle, viscosity):
    """echo_init_glauber_mhd on a resistive grid (B at the resistive slots,
    E and psi zero), IMEX-SSP3(4,3,3): against the oracle, then three
    steps."""
    from echoamd import EchoGrid
    rng = np.random.default_rng(12)
    ic, rc = make_config(nx=32, ny=28, nz=12, coordinates=2, mhd=2, divclean=1, algo_solver=7, nrk=4,
                         viscosity=viscosity, eqstate=1, xlim=(-6.0, 6.0), ylim=(-5.25, 5.25),
                         zlim=(-3.0, 3.0), tmin=0.6, init_type=3, maxdt=0.05, eta_over_s=0.1,
                         sigma_el=0.0058, sigma_chi=0.0015)
    pa = rng.normal(0, 2.0, (100, 2)) + [1.0, 0.0]
    pb = rng.normal(0, 2.0, (90, 2)) - [1.0, 0.0]
    pc = rng.normal(0, 1.6, (250, 2))
    nuc = _nucleons(rng)
    bp = dict(gamma_col=106.6, b_amplify=0.1, pw=0.5, magfield_type=1)
    print(f"\nglauber resistive: viscosity={viscosity}", flush=True)
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, PAR, pa, pb, pc)
    assert e == 0
    assert oracle_glauber_bfield(oracle, bp, nuc, v0) == 0
    assert oracle.init_from_prim(v0) == 0
    vo, uo = oracle.state()
    g = EchoGrid(ic, rc)
    assert g.init_glauber_mhd(PAR, pa, pb, pc, bp, nuc) == 0
    vg, ug = g.state()
    assert rel_linf(vg, vo) <= 1e-12
    assert rel_linf(ug, uo) <= 1e-12
    assert oracle.steps(3) == 0
    g.steps(3)
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= 1e-10 and rel_linf(ug, uo) <= 1e-10
    g.close()
