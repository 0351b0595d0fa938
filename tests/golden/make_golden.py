"""Writes the golden fixtures of tests/golden/ from the reference Fortran
(oracle/_ref/libechoref.so, built from the reference sources by
oracle/build_ref.sh).  Run in a container that has the reference:

    python tests/golden/make_golden.py

Each fixture holds inputs and the reference's outputs; the oracle is checked
against them by tests/test_oracle_golden.py where the reference library is
not available.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from libs import REC_NGC, Ref, make_config  # noqa: E402
from scenarios import fireball3d, gubser2d, shock2d  # noqa: E402


def rec_fixture(ref, rng):
    out = {}
    for code, ngc in REC_NGC.items():
        f = np.concatenate([np.sin(0.37 * np.arange(13 + 2 * ngc)),
                            rng.standard_normal(13 + 2 * ngc)])
        f = f.reshape(2, -1)
        gl, gr = zip(*(ref.rec(code, row) for row in f))
        out[f"rec{code}_f"] = f
        out[f"rec{code}_gl"] = np.array(gl)
        out[f"rec{code}_gr"] = np.array(gr)
    return out


def eos_fixture(ref, rng):
    out = {}
    en = np.exp(rng.uniform(np.log(1e-4), np.log(3000.0), 64))
    pr = en / 3.0 * rng.uniform(0.5, 1.0, en.size)
    rh = 1.0 + rng.uniform(size=en.size)
    for eq in (1, 2, 3):
        ic, rc = make_config(eqstate=eq)
        ref.setup(ic, rc)
        P = np.array([ref.eos_pressure(a, b)[1] for a, b in zip(rh, en)])
        E = np.array([ref.eos_energy(a, b)[1] for a, b in zip(rh, pr)])
        C = np.array([ref.eos_sound(a, b, c)[1] for a, b, c in zip(rh, en, pr)])
        D = np.array([ref.eos_derived(a, b)[1] for a, b in zip(rh, pr)])
        out.update({f"eos{eq}_pressure": P, f"eos{eq}_energy": E, f"eos{eq}_sound": C,
                    f"eos{eq}_derived": D, f"eos{eq}_przero": ref.c_przero()})
    out.update(eos_rh=rh, eos_en=en, eos_pr=pr)
    return out


def system_fixture(ref, rng):
    out = {}
    for visc in (0, 1):
        ic, rc = make_config(viscosity=visc, eqstate=3, coordinates=2)
        ref.setup(ic, rc)
        ref.set_time(0.6)
        nv = ref.nv
        n = 24
        v = np.zeros((n, nv))
        v[:, 0] = rng.uniform(0.5, 2.0, n)
        v[:, 1:4] = rng.uniform(-0.5, 0.5, (n, 3)) / np.sqrt(3) * np.array([1, 1, 1 / 0.6])
        v[:, 4] = np.exp(rng.uniform(np.log(0.01), np.log(50.0), n))
        if visc:
            v[:, 5:12] = rng.uniform(-0.05, 0.05, (n, 7)) * v[:, 4:5]
        U, Vp, F, VF, C2P = [], [], [], [], []
        for row in v:
            e, u, vout = ref.prim2cons(row)
            U.append(u)
            Vp.append(vout)
            F.append([ref.flux(row, d)[2] for d in (1, 2, 3)])
            VF.append([ref.flux(row, d)[3] for d in (1, 2, 3)])
            guess = row.copy()
            guess[4] *= 1.1
            C2P.append(ref.cons2prim(u, guess)[1])
        out.update({f"sys{visc}_v": v, f"sys{visc}_u": np.array(U), f"sys{visc}_vout": np.array(Vp),
                    f"sys{visc}_f": np.array(F), f"sys{visc}_vf": np.array(VF),
                    f"sys{visc}_c2p": np.array(C2P)})
    return out


TRAJ = {
    "gubser16_weno5": (lambda: gubser2d(16), 6),
    "shock16_mpe5": (lambda: shock2d(16, rec="MPE5"), 8),
    "fireball_visc_tab": (lambda: fireball3d(8, 6, 6, viscosity=1, eqstate=3), 4),
    "fireball_ideal_tab_rk2_mink": (lambda: fireball3d(8, 6, 6, eqstate=3, nrk=2, coordinates=1,
                                                       bc=0), 4),
}


def repair_fixture(ref):
    """The repair cases of tests/test_oracle_repair.py: inputs (initial
    primitives, and the host-forced state where there is one) and the
    reference's results."""
    from test_oracle_repair import FORCED, ROUGH, forced_state, u2v_state
    from scenarios import fast_eta
    out = {}
    cases = {}
    for name, (make, steps, _) in ROUGH.items():
        ic, rc, v0 = make()
        cases[name] = (ic, rc, v0, None, steps)
    for name, (kw, steps, _) in FORCED.items():
        ic, rc, v0, v, u = forced_state(ref, kw)
        cases[name] = (ic, rc, v0, (v, u), steps)
    ic, rc, v0 = fast_eta(rec="TVD2")
    cases["fast-eta-crash"] = (ic, rc, v0, None, 2)
    for name, (ic, rc, v0, st, steps) in cases.items():
        ref.setup(ic, rc)
        assert ref.init_from_prim(v0) == 0
        if st is not None:
            ref.set_state(*st)
            out[f"{name}/v_in"], out[f"{name}/u_in"] = st
        e = ref.steps(steps)
        v, u = ref.state()
        out.update({f"{name}/ic": ic, f"{name}/rc": rc, f"{name}/v0": v0, f"{name}/steps": steps,
                    f"{name}/err": e, f"{name}/v": v, f"{name}/u": u})
    # evolve_u2v alone
    from libs import Oracle
    ic, rc, v0, v, u = u2v_state(Oracle())
    ref.setup(ic, rc)
    ref.init_from_prim(v0)
    ref.set_state(v, u)
    e = ref.u2v()
    vr, ur = ref.state()
    out.update({"u2v/ic": ic, "u2v/rc": rc, "u2v/v0": v0, "u2v/v_in": v, "u2v/u_in": u, "u2v/err": e,
                "u2v/v": vr, "u2v/u": ur})
    return out


def main():
    ref = Ref()
    rng = np.random.default_rng(42)
    np.savez_compressed(os.path.join(HERE, "rec.npz"), **rec_fixture(ref, rng))
    np.savez_compressed(os.path.join(HERE, "eos.npz"), **eos_fixture(ref, rng))
    np.savez_compressed(os.path.join(HERE, "system.npz"), **system_fixture(ref, rng))
    np.savez_compressed(os.path.join(HERE, "repair.npz"), **repair_fixture(ref))
    for name, (make, n) in TRAJ.items():
        ic, rc, v0 = make()
        ref.setup(ic, rc)
        assert ref.init_from_prim(v0) == 0
        assert ref.steps(n) == 0
        v, u = ref.state()
        s = ref.scalars()
        np.savez_compressed(os.path.join(HERE, f"traj_{name}.npz"), ic=ic, rc=rc, v0=v0,
                            nsteps=n, v=v, u=u, t=s["t"], dt=s["dt"])
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
