"""The Fortran caller (echo-qgp-mhd-devel_amd/fortran/echo_amd_driver, built
with flang on the iso_c_binding module echo_amd_f) runs a param.dat-format
Gubser case on the GPU; its final state equals the oracle's to 1e-10."""
import os
import subprocess

import numpy as np
import pytest

from libs import PKG, Oracle
from scenarios import GUBSER_THAT0, gubser2d, rel_linf

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(PKG, "fortran", "echo_amd_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_gubser(oracle, tmp_path):
    n, steps = 32, 12
    ic, rc, v0 = gubser2d(n=n)
    half = n * 0.1 / 2
    dump = tmp_path / "state.bin"
    lines = ["! Gubser flow, param.dat format", "INIT_TYPE=4", "COORD....=2",
             f"NX.......={n}", f"NY.......={n}", "NZ.......=1",
             f"XMIN.....={-half!r}", f"XMAX.....={half!r}", f"YMIN.....={-half!r}",
             f"YMAX.....={half!r}", "ZMIN.....=-0.5", "ZMAX.....=0.5", "TSTART...=1.0",
             "CFL......=0.4", "MAXDT....=1.0", "REC_ALGO.=WENO5", "INT_ALGO.=RK3",
             "EOS......=1", "bound_con=2", f"GUB_T0...={GUBSER_THAT0!r}",
             f"NSTEPS...={steps}", "CHUNK....=5", f"DUMP_FILE={dump}"]
    pf = tmp_path / "param.dat"
    pf.write_text("\n".join(lines) + "\n")
    r = subprocess.run([DRIVER, str(pf)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:])
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(dump, dtype=np.float64)
    nv = 5
    v = raw[:nv * n * n].reshape(nv, 1, n, n)
    u = raw[nv * n * n:].reshape(nv, 1, n, n)
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(steps) == 0
    vo, uo = oracle.state()
    assert rel_linf(v, vo) <= 1e-10
    assert rel_linf(u, uo) <= 1e-10


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_freezeout(oracle, tmp_path):
    """HYP_COMPU=1: the driver writes hypersurface.txt at the reference's
    cadence (echo.f90:252-258); the oracle, run through the same cadence,
    writes the same file to the 8 digits of its format."""
    from libs import OracleFO
    n, steps, fv, fdt = 32, 12, 0.7, 0.02
    ic, rc, v0 = gubser2d(n=n)
    half = n * 0.1 / 2
    hyp = tmp_path / "hyp_gpu.txt"
    lines = ["INIT_TYPE=4", "COORD....=2", f"NX.......={n}", f"NY.......={n}", "NZ.......=1",
             f"XMIN.....={-half!r}", f"XMAX.....={half!r}", f"YMIN.....={-half!r}",
             f"YMAX.....={half!r}", "ZMIN.....=-0.5", "ZMAX.....=0.5", "TSTART...=1.0",
             "CFL......=0.4", "MAXDT....=1.0", "REC_ALGO.=WENO5", "INT_ALGO.=RK3",
             "EOS......=1", "bound_con=2", f"GUB_T0...={GUBSER_THAT0!r}",
             f"NSTEPS...={steps}", "HYP_COMPU=1", "FREEZKIND=0", f"FREEZEVAL={fv!r}",
             f"HYPSURFTI={fdt!r}", f"HYP_FILE.={hyp}"]
    pf = tmp_path / "param.dat"
    pf.write_text("\n".join(lines) + "\n")
    r = subprocess.run([DRIVER, str(pf)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:])
    assert r.returncode == 0, r.stdout + r.stderr
    oracle.setup(ic, rc)
    assert oracle.init_from_prim(v0) == 0
    fo = OracleFO(oracle)
    assert fo.setup(0, fv)[0] == 0
    opath = str(tmp_path / "hyp_orc.txt")
    assert fo.write(opath, 1) == 0
    thyp, tmax, first = rc[6], 1e6, True
    for _ in range(steps):
        assert oracle.steps(1) == 0
        t = oracle.scalars()["t"]
        if t > thyp or thyp - t < 1e-10:
            thyp = min(thyp + fdt, tmax)
            assert fo.slice()[0] == 0
            if not first:
                assert fo.write(opath, 0) == 0
            first = False
    a = open(hyp).read().splitlines()
    b = open(opath).read().splitlines()
    assert len(a) == len(b) > 100
    assert a[0].split()[:2] == b[0].split()[:2]
    for la, lb in zip(a[1:], b[1:]):
        va, vb = np.array(la.split(), dtype=float), np.array(lb.split(), dtype=float)
        assert np.all(np.abs(va - vb) <= 1.5e-7 * np.abs(vb) + 1e-300), (la, lb)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_restart(oracle, tmp_path):
    """RESTART..=1 writes dump.dat at the output times; RESUME...= starts a
    second run from it.  The oracle, restarted from the same file, takes the
    same steps to 1e-10."""
    from libs import oracle_restart
    n, dtout = 32, 0.05
    ic, rc, v0 = gubser2d(n=n)
    half = n * 0.1 / 2
    base = ["INIT_TYPE=4", "COORD....=2", f"NX.......={n}", f"NY.......={n}", "NZ.......=1",
            f"XMIN.....={-half!r}", f"XMAX.....={half!r}", f"YMIN.....={-half!r}",
            f"YMAX.....={half!r}", "ZMIN.....=-0.5", "ZMAX.....=0.5", "TSTART...=1.0",
            "CFL......=0.4", "MAXDT....=1.0", "REC_ALGO.=WENO5", "INT_ALGO.=RK3",
            "EOS......=1", "bound_con=2", f"GUB_T0...={GUBSER_THAT0!r}", f"DTOUT....={dtout!r}"]
    run1 = tmp_path / "run1.dat"
    run1.write_text("\n".join(base + ["NSTEPS...=6", "RESTART..=1", f"DUMP_DIR.={tmp_path}"]) + "\n")
    r = subprocess.run([DRIVER, str(run1)], cwd=PKG, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dump = tmp_path / "dump.dat"
    assert dump.exists()
    final = tmp_path / "final.bin"
    run2 = tmp_path / "run2.dat"
    run2.write_text("\n".join(base + ["NSTEPS...=5", f"RESUME...={dump}", f"DUMP_FILE={final}"]) + "\n")
    r = subprocess.run([DRIVER, str(run2)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "restarted from" in r.stdout
    raw = np.fromfile(final, dtype=np.float64)
    nv = 5
    v = raw[:nv * n * n].reshape(nv, 1, n, n)
    u = raw[nv * n * n:].reshape(nv, 1, n, n)
    rc2 = rc.copy()
    rc2[16] = dtout
    oracle.setup(ic, rc2)
    e, iout, _ = oracle_restart(oracle, dump)
    assert e == 0 and iout >= 0
    assert oracle.steps(5) == 0
    vo, uo = oracle.state()
    assert rel_linf(v, vo) <= 1e-10
    assert rel_linf(u, uo) <= 1e-10


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_glauber(oracle, tmp_path):
    """INIT_TYPE=3: the driver generates Au-Au events (echo_glauber_generate,
    SEED 9, b = 6 fm fixed), reads event 1 back, deposits it on the device
    and runs viscous hydro with the tabulated EoS; it writes frames of the
    selected fields at the start and at the output times.  The final state
    equals the oracle's run from the same event (read in Python with
    echo_glauber_read_event) to 1e-10.  A second run with TEMP_END above
    the largest energy density stops before its first step and writes its
    last frame."""
    from echoamd import glauber_read_event
    from libs import make_config, oracle_glauber_prim
    n, nz, steps = 40, 8, 4
    gdir = tmp_path / "events"
    gdir.mkdir()
    out = tmp_path / "frames"
    out.mkdir()
    dump = tmp_path / "state.bin"
    base = ["INIT_TYPE=3", "COORD....=2", "VISCOUS..=1", "EOS......=3", "EOS_FILE.=qcdIEOS0.dat",
            f"NX.......={n}", f"NY.......={n}", f"NZ.......={nz}", "XMIN.....=-8.0", "XMAX.....=8.0",
            "YMIN.....=-8.0", "YMAX.....=8.0", "ZMIN.....=-2.0", "ZMAX.....=2.0", "TSTART...=0.6",
            "CFL......=0.4", "MAXDT....=0.05", "REC_ALGO.=WENO5", "INT_ALGO.=RK3", "bound_con=2",
            "ETA_S....=0.1", "NUCLEUS..=Au", "SIGMA_IN.=42.", "B........=6.0", "FIXED_B..=1", "NCONF....=1",
            "KAPPA....=19.", "AH.......=0.15", "SIG......=0.4", "DETA.....=2.0", "SIGETA...=0.8",
            "RHOCENTER=1.", "RADS.....=200.", "SEED.....=9", f"GLAUB_DIR={gdir}", "DTOUT....=0.04",
            f"outdir...={out}", "density..=1", "pressure.=1", "temper...=1", "OUTP_PREC=8"]
    pf = tmp_path / "param.dat"
    pf.write_text("\n".join(base + [f"NSTEPS...={steps}", "CHUNK....=1", f"DUMP_FILE={dump}"]) + "\n")
    r = subprocess.run([DRIVER, str(pf)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Glauber event" in r.stdout
    ev = glauber_read_event(gdir, 1)
    assert abs(ev["b"] - 6.0) < 1e-12
    ic, rc = make_config(nx=n, ny=n, nz=nz, coordinates=2, viscosity=1, eqstate=3, rec="WENO5", nrk=3, bc=2,
                         init_type=3, xlim=(-8.0, 8.0), ylim=(-8.0, 8.0), zlim=(-2.0, 2.0), tmin=0.6,
                         cfl=0.4, maxdt=0.05, eta_over_s=0.1, dtout=0.04)
    par = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=np.log(200.0 / 0.93891897), deta=2.0, sigeta=0.8,
               rhocenter=1.0)
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, par, ev["part_a"], ev["part_b"], ev["coll"])
    assert e == 0 and oracle.init_from_prim(v0) == 0
    assert oracle.steps(steps) == 0
    vo, uo = oracle.state()
    raw = np.fromfile(dump, dtype=np.float64)
    m = 12 * n * n * nz
    v = raw[:m].reshape(12, nz, n, n)
    u = raw[m:].reshape(12, nz, n, n)
    assert rel_linf(v, vo) <= 1e-10 and rel_linf(u, uo) <= 1e-10
    frames = sorted(p.name for p in out.glob("out*.dat"))
    assert frames[0] == "out0000.dat" and len(frames) >= 2
    # t, then density, pressure, temperature: 3 fields of n*n*nz FP64
    assert (out / "out0000.dat").stat().st_size == 8 * (1 + 3 * n * n * nz)
    # TEMP_END above every cell's energy density (FREEZKIND 1): no step
    out2 = tmp_path / "frames2"
    out2.mkdir()
    pf2 = tmp_path / "param2.dat"
    pf2.write_text("\n".join([x for x in base if not x.startswith("outdir")]
                             + [f"outdir...={out2}", "TEMP_END.=1.e6", "NSTEPS...=3"]) + "\n")
    r = subprocess.run([DRIVER, str(pf2)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Energy density is under" in r.stdout
    assert (out2 / "out0001.dat").exists()


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_glauber_resistive(oracle, tmp_path):
    """The driver on the viscous resistive system (MHD=2, VISCOUS=1,
    INT_ALGO=SSP, tabulated EoS) from a generated Au-Au event with the field
    of its protons (MAGFIELDT 1, B_amp_fac 1, PW 0.5: the reference's
    defaults): frames with B and E, a restart dump at the output time, and a
    final state equal to the oracle's run from the same event to 1e-10."""
    from echoamd import glauber_read_event
    from libs import make_config, oracle_glauber_bfield, oracle_glauber_prim
    n, nz, steps = 40, 8, 4
    gdir = tmp_path / "events"
    gdir.mkdir()
    out = tmp_path / "frames"
    out.mkdir()
    dump = tmp_path / "state.bin"
    lines = ["INIT_TYPE=3", "COORD....=2", "VISCOUS..=1", "MHD......=2", "DIVCLEAN.=1", "ALG_SOLV.=7",
             "EOS......=3", "EOS_FILE.=qcdIEOS0.dat", f"NX.......={n}", f"NY.......={n}", f"NZ.......={nz}",
             "XMIN.....=-8.0", "XMAX.....=8.0", "YMIN.....=-8.0", "YMAX.....=8.0", "ZMIN.....=-2.0",
             "ZMAX.....=2.0", "TSTART...=0.6", "CFL......=0.4", "MAXDT....=0.05", "REC_ALGO.=WENO5",
             "INT_ALGO.=SSP", "bound_con=2", "ETA_S....=0.1", "NUCLEUS..=Au", "SIGMA_IN.=42.", "B........=6.0",
             "FIXED_B..=1", "NCONF....=1", "KAPPA....=19.", "AH.......=0.15", "SIG......=0.4", "DETA.....=2.0",
             "SIGETA...=0.8", "RHOCENTER=1.", "RADS.....=200.", "SEED.....=9", f"GLAUB_DIR={gdir}",
             "MAGFIELDT=1", "B_amp_fac=1.", "PW.......=0.5", "DTOUT....=0.1", "RESTART..=1",
             f"DUMP_DIR.={tmp_path}", f"outdir...={out}", "density..=1", "bx.......=1", "ex.......=1",
             "OUTP_PREC=8", f"NSTEPS...={steps}", "CHUNK....=1", f"DUMP_FILE={dump}"]
    pf = tmp_path / "param.dat"
    pf.write_text("\n".join(lines) + "\n")
    print("\nfortran driver: viscous resistive IMEX from a Glauber event", flush=True)
    r = subprocess.run([DRIVER, str(pf)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:])
    assert r.returncode == 0, r.stdout + r.stderr
    ev = glauber_read_event(gdir, 1)
    ic, rc = make_config(nx=n, ny=n, nz=nz, coordinates=2, viscosity=1, eqstate=3, rec="WENO5", nrk=4, bc=2,
                         mhd=2, divclean=1, algo_solver=7, init_type=3, xlim=(-8.0, 8.0), ylim=(-8.0, 8.0),
                         zlim=(-2.0, 2.0), tmin=0.6, cfl=0.4, maxdt=0.05, eta_over_s=0.1, dtout=0.1)
    par = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=np.log(200.0 / 0.93891897), deta=2.0, sigeta=0.8,
               rhocenter=1.0)
    bp = dict(gamma_col=200.0 / (2 * 0.93891897), b_amplify=1.0, pw=0.5, magfield_type=1)
    lists = [(ev["spec_a"], ev["type_spec_a"]), (ev["part_a"], ev["type_part_a"]),
             (ev["spec_b"], ev["type_spec_b"]), (ev["part_b"], ev["type_part_b"])]
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, par, ev["part_a"], ev["part_b"], ev["coll"])
    assert e == 0
    assert oracle_glauber_bfield(oracle, bp, lists, v0) == 0
    assert np.abs(v0[12:15]).max() > 0.0
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(steps) == 0
    vo, uo = oracle.state()
    raw = np.fromfile(dump, dtype=np.float64)
    m = 19 * n * n * nz
    v = raw[:m].reshape(19, nz, n, n)
    u = raw[m:].reshape(19, nz, n, n)
    assert rel_linf(v, vo) <= 1e-10 and rel_linf(u, uo) <= 1e-10
    # frames: t, density, B^x, E^x
    f0 = np.fromfile(out / "out0000.dat", dtype=np.float64)
    assert f0.size == 1 + 3 * n * n * nz
    # device and glibc atan2, sin, cos, exp, pow differ in the last bits
    bx = f0[1 + n * n * nz:1 + 2 * n * n * nz]
    assert np.abs(bx - v0[12].ravel()).max() <= 1e-12 * np.abs(v0[12]).max()
    assert not f0[1 + 2 * n * n * nz:].any()          # E starts at zero
    assert (tmp_path / "dump.dat").exists()


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_driver_glauber_ideal_mhd_imex(oracle, tmp_path):
    """The driver on ideal MHD (MHD=1, GLM, p = e/3 with algorithm 5) with
    INT_ALGO=SSP from a generated Au-Au event and its protons' field: the
    final state equals the oracle's to 1e-10.  The field is scaled by 0.1
    (B_amp_fac): at the reference's default the near-vacuum corners of this
    small grid reach B^2/p ~ 30, where the inviscid inversion amplifies
    rounding beyond the bar."""
    from echoamd import glauber_read_event
    from libs import make_config, oracle_glauber_bfield, oracle_glauber_prim
    n, nz, steps = 32, 8, 4
    gdir = tmp_path / "events"
    gdir.mkdir()
    dump = tmp_path / "state.bin"
    lines = ["INIT_TYPE=3", "COORD....=2", "MHD......=1", "DIVCLEAN.=1", "ALG_SOLV.=5", "EOS......=1",
             f"NX.......={n}", f"NY.......={n}", f"NZ.......={nz}", "XMIN.....=-8.0", "XMAX.....=8.0",
             "YMIN.....=-8.0", "YMAX.....=8.0", "ZMIN.....=-2.0", "ZMAX.....=2.0", "TSTART...=0.6",
             "CFL......=0.4", "MAXDT....=0.05", "REC_ALGO.=WENO5", "INT_ALGO.=SSP", "bound_con=2",
             "NUCLEUS..=Au", "SIGMA_IN.=42.", "B........=6.0", "FIXED_B..=1", "NCONF....=1", "KAPPA....=19.",
             "AH.......=0.15", "SIG......=0.4", "DETA.....=2.0", "SIGETA...=0.8", "RHOCENTER=1.",
             "RADS.....=200.", "SEED.....=9", f"GLAUB_DIR={gdir}", "MAGFIELDT=1", "B_amp_fac=0.1",
             "PW.......=0.5", "DTOUT....=1.0", f"NSTEPS...={steps}", "CHUNK....=2", f"DUMP_FILE={dump}"]
    pf = tmp_path / "param.dat"
    pf.write_text("\n".join(lines) + "\n")
    print("\nfortran driver: ideal MHD IMEX from a Glauber event", flush=True)
    r = subprocess.run([DRIVER, str(pf)], cwd=PKG, capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stdout + r.stderr
    ev = glauber_read_event(gdir, 1)
    ic, rc = make_config(nx=n, ny=n, nz=nz, coordinates=2, mhd=1, divclean=1, algo_solver=5, eqstate=1,
                         rec="WENO5", nrk=4, bc=2, init_type=3, xlim=(-8.0, 8.0), ylim=(-8.0, 8.0),
                         zlim=(-2.0, 2.0), tmin=0.6, cfl=0.4, maxdt=0.05, dtout=1.0)
    par = dict(kappa=19.0, ah=0.15, sig_mc=0.4, ybeam=np.log(200.0 / 0.93891897), deta=2.0, sigeta=0.8,
               rhocenter=1.0)
    bp = dict(gamma_col=200.0 / (2 * 0.93891897), b_amplify=0.1, pw=0.5, magfield_type=1)
    lists = [(ev["spec_a"], ev["type_spec_a"]), (ev["part_a"], ev["type_part_a"]),
             (ev["spec_b"], ev["type_spec_b"]), (ev["part_b"], ev["type_part_b"])]
    oracle.setup(ic, rc)
    e, v0 = oracle_glauber_prim(oracle, par, ev["part_a"], ev["part_b"], ev["coll"])
    assert e == 0
    assert oracle_glauber_bfield(oracle, bp, lists, v0) == 0
    assert oracle.init_from_prim(v0) == 0
    assert oracle.steps(steps) == 0
    vo, uo = oracle.state()
    raw = np.fromfile(dump, dtype=np.float64)
    m = 9 * n * n * nz
    v = raw[:m].reshape(9, nz, n, n)
    u = raw[m:].reshape(9, nz, n, n)
    assert rel_linf(v, vo) <= 1e-10 and rel_linf(u, uo) <= 1e-10
