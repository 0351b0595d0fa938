"""Data, initialisation and training driver of the learned con2prim.

The inversion itself (network on MFMA, FP64 Newton refinement, residual gate,
plain-solver fallback) and the trainer's kernels live in libechoamd.so
(``C2pNet``); this module prepares their inputs on the host.

States: ideal MHD with p = e/3 in the local orthonormal (Minkowski) frame,
conserved (D, S_x, S_y, S_eta, U, B^x, B^y, B^eta) as system_prim2cons forms
them (system.f90:164-263; sqrt(-g) = 1).  They are drawn around snapshot
cells of the Gubser-flow run: each state takes the snapshot's (rho, p) and
flow direction, a Lorentz factor log-uniform in [1, 20] and a magnetisation
sigma = B^2/(e + p) log-uniform in [1e-4, 1] with a random field direction
(the "base" state, which is the warm-start guess); the target state then
perturbs rho, p, gamma*v, |B| by log-uniform factors in [1/1.1, 1.1] and tilts
both directions by 10 %.
"""
import os

import numpy as np

from . import MLP_NPARAM, PKG_DIR

WEIGHTS = os.path.join(PKG_DIR, "weights", "c2p_e3p_mlp.npy")

_DIN = (8, 64, 64, 64)
_DOUT = (64, 64, 64, 5)


def init_params(seed=0):
    """Uniform +-1/sqrt(fan_in) weights, zero biases, in the C-ABI layout."""
    rng = np.random.default_rng(seed)
    out = []
    for ni, no in zip(_DIN, _DOUT):
        out.append(rng.uniform(-1.0, 1.0, (no, ni)).ravel() / np.sqrt(ni))
        out.append(np.zeros(no))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == MLP_NPARAM
    return p


def load_weights(path=WEIGHTS):
    p = np.load(path, allow_pickle=False)
    assert p.dtype == np.float32 and p.size == MLP_NPARAM, (p.dtype, p.size)
    return p


def _unit(rng, n):
    g = rng.standard_normal((3, n))
    return g / np.linalg.norm(g, axis=0)


def _tilt(rng, d, amount):
    t = d + amount * rng.standard_normal(d.shape)
    return t / np.linalg.norm(t, axis=0)


def prim2cons_e3p(v):
    """system_prim2cons for ideal MHD, p = e/3, Minkowski (system.f90:164-263
    with compute_E_ideal and system_EXB, system.f90:1846-1862).
    v: (8, n) = (rho, v^i, p, B^i).  Returns u (8, n)."""
    rho, vel, p, B = v[0], v[1:4], v[4], v[5:8]
    v2 = (vel * vel).sum(0)
    glf2 = 1.0 / (1.0 - v2)
    ww = 4.0 * p * glf2
    E = -np.cross(vel, B, axis=0)
    exb = np.cross(E, B, axis=0)
    E2 = (E * E).sum(0)
    B2 = (B * B).sum(0)
    u = np.empty_like(v)
    u[0] = rho * np.sqrt(glf2)
    u[1:4] = ww * vel + exb
    u[4] = ww - p + 0.5 * (E2 + B2)
    u[5:8] = B
    return u


def mhd_e3p_states(n, snap, seed=0, gamma_max=20.0, sigma_range=(1e-4, 1.0), pert=0.1):
    """n seeded states around snapshot cells.

    snap: (5, m) primitives (rho, v^x, v^y, v^eta (orthonormal), p) of the
    snapshot cells.  Returns (u, v_true, v_warm), each (8, n): conserved
    states, the primitives they were built from, and the unperturbed base
    primitives (the warm-start guess)."""
    rng = np.random.default_rng(seed)
    snap = np.asarray(snap, dtype=np.float64)
    c = rng.integers(0, snap.shape[1], n)
    rho0, p0 = snap[0, c], snap[4, c]
    vs = snap[1:4, c]
    sp = np.linalg.norm(vs, axis=0)
    d0 = np.where(sp > 1e-12, vs / np.maximum(sp, 1e-300), _unit(rng, n))
    gam0 = np.exp(rng.uniform(0.0, np.log(gamma_max), n))
    sig0 = np.exp(rng.uniform(np.log(sigma_range[0]), np.log(sigma_range[1]), n))
    b0 = _unit(rng, n)
    u0 = np.sqrt(gam0 ** 2 - 1.0)                      # |gamma v|
    warm = np.empty((8, n))
    warm[0] = rho0
    warm[1:4] = d0 * (u0 / gam0)
    warm[4] = p0
    warm[5:8] = b0 * np.sqrt(sig0 * 4.0 * p0)

    def fac():
        return np.exp(rng.uniform(-np.log(1.0 + pert), np.log(1.0 + pert), n))

    rho1, p1, u1 = rho0 * fac(), p0 * fac(), u0 * fac()
    d1 = _tilt(rng, d0, pert)
    b1 = _tilt(rng, b0, pert)
    g1 = np.sqrt(1.0 + u1 ** 2)
    true = np.empty((8, n))
    true[0] = rho1
    true[1:4] = d1 * (u1 / g1)
    true[4] = p1
    true[5:8] = b1 * np.sqrt(sig0 * 4.0 * p0) * fac()
    return prim2cons_e3p(true), true, warm


def gubser_snapshots(grid, steps=(0, 20, 40, 60, 80, 100)):
    """Primitive snapshots (rho, v^x, v^y, tau v^eta, p) of a running Gubser
    grid (an EchoGrid of the 128 x 128 x 1 Bjorken run) at the given step
    counts; returns (5, m)."""
    out = []
    done = 0
    for k in sorted(steps):
        if k > done:
            grid.steps(k - done)
            done = k
        v, _ = grid.state()
        t = grid.scalars()["t"]
        s = np.stack([v[0], v[1], v[2], t * v[3], v[4]]).reshape(5, -1)
        out.append(s)
    return np.concatenate(out, axis=1)


def gubser_grid(device=0, n=128, dx=0.1):
    """The config-1 Gubser-flow grid (128 x 128 x 1 Bjorken cells of 0.1 fm,
    p = e/3, WENO5 + HLL + RK3, CFL 0.4, T0 = 1.2 GeV at r = 0, tau0 = 1 fm),
    initialised; an EchoGrid."""
    from . import EchoGrid
    from .initial import HBAR, gubser_prim, make_config
    half = n * dx / 2
    ic, rc = make_config(nx=n, ny=n, nz=1, coordinates=2, rec="WENO5", nrk=3, bc=2, eqstate=1,
                         init_type=4, xlim=(-half, half), ylim=(-half, half), zlim=(-0.5, 0.5),
                         tmin=1.0, cfl=0.4, maxdt=1.0)
    g = EchoGrid(ic, rc, device=device)
    g.init_from_prim(gubser_prim(n, n, (-half, half), (-half, half), 1.0, that0=1.2 / HBAR))
    return g


def e3p_grid(device=0, coordinates=1, tmin=1.0, divclean=0):
    """A small ideal-MHD p = e/3 grid (algorithm 5) whose metric, stream and
    plain solver the batched inversions use."""
    from . import EchoGrid
    from .initial import make_config
    ic, rc = make_config(nx=8, ny=8, nz=1, coordinates=coordinates, mhd=1, divclean=divclean,
                         eqstate=1, algo_solver=5, tmin=tmin)
    return EchoGrid(ic, rc, device=device)


def plain_labels(grid, u, guess):
    """Plain-solver primitives of host states u (8, n) from the guess; returns
    (v, status) on the host."""
    from . import DeviceArray
    du = DeviceArray.from_numpy(u)
    dv = DeviceArray.from_numpy(guess)
    ds = DeviceArray((u.shape[1],), np.int32)
    grid.cons2prim_batch_dev(du.ptr, dv.ptr, ds.ptr, u.shape[1])
    return dv.numpy(), ds.numpy()


def train(device=0, n_train=1 << 20, steps=20000, batch=8192, lr=3e-3, lr_final=3e-5, seed=1,
          params=None, snapshots=None, log=print):
    """Train the network on the device from plain-solver labels.

    Returns (params, loss_log).  The training states come from
    mhd_e3p_states around Gubser snapshots; their labels are the primitives
    the plain FP64 solver (algorithm 5, warm-started) returns."""
    from . import C2pNet, DeviceArray
    if snapshots is None:
        g1 = gubser_grid(device)
        snapshots = gubser_snapshots(g1)
        g1.close()
    grid = e3p_grid(device)
    u, _, warm = mhd_e3p_states(n_train, snapshots, seed=seed)
    v, st = plain_labels(grid, u, warm)
    ok = st == 0
    log(f"train: {ok.sum()} of {n_train} labelled states converged in the plain solver")
    u = np.ascontiguousarray(u[:, ok])
    v = np.ascontiguousarray(v[:, ok])
    n = u.shape[1]
    du = DeviceArray.from_numpy(u)
    dv = DeviceArray.from_numpy(v)
    net = C2pNet(device, params if params is not None else init_params(seed))
    loss = net.train(grid, du.ptr, dv.ptr, n, steps, min(batch, n), lr, lr_final)
    p = net.params()
    net.close()
    grid.close()
    return p, loss


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train the learned con2prim network on GPU 0")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--steps", type=int, default=60000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--n-train", type=int, default=1 << 20)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    import time
    t0 = time.perf_counter()
    p, loss = train(n_train=a.n_train, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed)
    print(f"trained {a.steps} steps of {a.batch} states in {time.perf_counter() - t0:.1f} s "
          "(data generation and labelling included)")
    k = max(1, len(loss) // 20)
    for i in range(0, len(loss), k):
        print(f"step {i:6d} loss {loss[i:i + k].mean():.3e}", flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    np.save(a.out, p)
    print(f"saved {a.out}")


if __name__ == "__main__":
    main()
