"""The learned-con2prim oracle (oracle/oracle_mlp.c) pinned by mathematics.

The reference has no learned inversion, so these checks fix each part of the
restatement independently of it:
- backward pass: central finite differences of the FP64 forward pass;
- Newton refinement + primitive formulas: convergence to the root that the
  reference's own algorithm 5 (restated in oracle_sys.c, bit-exact against
  the reference Fortran) reaches, and to the primitives the state was built
  from; the refinement's residual vanishes at the exact solution and its
  derivative matches finite differences;
- residual gate: rejects an unconverged guess, accepts a converged one;
- features: invariance under rotations and under the scaling
  (D, S, U, B) -> (l D, l S, l U, sqrt(l) B) that maps solutions of the
  p = e/3 inversion onto each other;
- the host-side data generator's prim2cons against the oracle's.
"""
import numpy as np
import pytest

from libs import OracleMLP, gubser_prim, make_config


@pytest.fixture(scope="module")
def om(oracle):
    return OracleMLP()


def _snap():
    v = gubser_prim(24, 24, (-2.4, 2.4), (-2.4, 2.4), 1.0)
    return np.stack([v[0, 0], v[1, 0], v[2, 0], 0 * v[3, 0], v[4, 0]]).reshape(5, -1)


def _states(n, seed=3):
    from echoamd.surrogate import mhd_e3p_states
    return mhd_e3p_states(n, _snap(), seed=seed)


def _scales(v):
    """Per-component error scales: rho and p their own size, velocity and
    field components the largest component of their vector."""
    return np.array([abs(v[0])] + [np.abs(v[1:4]).max()] * 3 + [abs(v[4])] +
                    [np.abs(v[5:8]).max()] * 3)


def _mhd_oracle(oracle, coords=1, t=1.0):
    ic, rc = make_config(mhd=1, divclean=0, algo_solver=5, eqstate=1, coordinates=coords)
    oracle.setup(ic, rc)
    oracle.set_time(t)


def test_f16_rounding(om):
    """IEEE binary16, nearest with ties to even: known cases and numpy's
    float16 conversion (an independent implementation) on random values
    over the normal, subnormal and overflow ranges."""
    assert om.f16(1.0) == 1.0
    assert om.f16(1.0 + 2 ** -11) == 1.0                  # tie -> even
    assert om.f16(1.0 + 3 * 2 ** -11) == 1.0 + 2 ** -9    # tie -> even (up)
    assert om.f16(1.0 + 2 ** -11 + 2 ** -30) == 1.0 + 2 ** -10
    assert om.f16(-3.0) == -3.0
    assert om.f16(2.0 ** -25) == 0.0                      # half the smallest subnormal
    assert om.f16(3 * 2.0 ** -26) == 2.0 ** -24
    assert om.f16(65504.0) == 65504.0
    assert om.f16(65519.9) == 65504.0
    assert om.f16(65520.0) == np.inf
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-8, 8, 20000),
                        np.exp(rng.uniform(np.log(1e-9), np.log(7e4), 20000)) * rng.choice([-1, 1], 20000)])
    x = x.astype(np.float32)
    mine = np.array([om.f16(float(a)) for a in x])
    ref = x.astype(np.float16).astype(np.float64)
    assert np.array_equal(mine, ref)


def test_half_precision_activation(om):
    """orc_act_h (the kernel's activation in half precision) against the
    same three operations in numpy float16: each product of two float16
    values and its sum with 1/2 are exact in float64, so rounding them once
    to float16 is the hardware's fused multiply-add and multiply."""
    c6 = np.float64(np.float16(np.float32(0.16666667)))
    rng = np.random.default_rng(4)
    z = np.concatenate([rng.uniform(-5, 5, 20000), rng.normal(0, 0.05, 5000),
                        [-3.0, 3.0, -1.5, 0.0, 1e-6, -1e-6, 100.0, -100.0]]).astype(np.float32)
    x = z.astype(np.float16).astype(np.float64)
    t = np.clip(np.float16(x * c6 + 0.5).astype(np.float64), 0.0, 1.0)
    ref = np.float16(x * t).astype(np.float64)
    mine = np.array([om.act_h(float(a)) for a in z])
    assert np.array_equal(mine, ref)
    # and it is the float activation to half-precision accuracy: rounding z,
    # t and h each to 11 bits moves h by at most ~3 ulps of max(|z|, 1)
    hf = np.array([om.act(float(a)) for a in z])
    assert np.all(np.abs(mine - hf) <= 4 * 2.0 ** -11 * np.maximum(np.abs(z), 1.0))


def test_activation_shape(om):
    """hard-swish: 0 below -3, the identity above 3, continuous at both
    joins, x^2/6 + x/2 between (to float rounding), monotone above its
    minimum -3/8 at x = -3/2; the derivative against central differences
    away from the joins"""
    for x in (-50.0, -7.5, -3.0001, -3.0):
        assert om.act(x) == 0.0 or abs(om.act(x)) < 1e-6
    for x in (3.0001, 4.0, 11.0, 1e4):
        assert abs(om.act(x) - x) <= 1e-6 * x
    x = np.linspace(-2.99, 2.99, 599)
    h = np.array([om.act(float(a)) for a in x])
    assert np.max(np.abs(h - (x * x / 6 + x / 2))) < 2e-6
    assert abs(min(h) + 3 / 8) < 1e-3
    xr = x[x > -1.5]
    hr = np.array([om.act(float(a)) for a in xr])
    assert np.all(np.diff(hr) >= 0)
    for a in np.concatenate([np.linspace(-2.9, 2.9, 37), [-4.0, 4.0]]):
        d = 1e-3
        fd = (om.act(a + d) - om.act(a - d)) / (2 * d)
        assert abs(om.act_d(a) - fd) < 2e-3, a


def test_backward_matches_finite_differences(om):
    from echoamd.surrogate import init_params
    u, vt, _ = _states(12)
    P = init_params(5)
    rng = np.random.default_rng(0)
    P[512:576] = rng.uniform(-0.2, 0.2, 64)     # non-zero biases
    P[4672:4736] = rng.uniform(-0.2, 0.2, 64)
    L, g = om.loss(P, u, vt, rnd=0, grad=True)
    assert np.isfinite(L) and L > 0
    idx = np.concatenate([rng.choice(np.arange(0, 512), 8, replace=False),
                          rng.choice(np.arange(512, 576), 4, replace=False),
                          rng.choice(np.arange(576, 4672), 8, replace=False),
                          rng.choice(np.arange(4736, 8832), 8, replace=False),
                          rng.choice(np.arange(8896, 9221), 8, replace=False)])
    h = 2.0 ** -12
    for i in idx:
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        fd = (om.loss(Pp, u, vt, rnd=0) - om.loss(Pm, u, vt, rnd=0)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g).max()) + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_f16_gradient_close_to_fp64_gradient(om):
    from echoamd.surrogate import init_params
    u, vt, _ = _states(64)
    P = init_params(6)
    L0, g0 = om.loss(P, u, vt, rnd=0, grad=True)
    L1, g1 = om.loss(P, u, vt, rnd=1, grad=True)
    assert abs(L1 - L0) < 2e-2 * L0
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)


def test_targets_encode_the_solution(om):
    u, vt, _ = _states(50)
    for s in range(50):
        t = om.targets(u[:, s], vt[:, s])
        v2 = (vt[1:4, s] ** 2).sum()
        assert abs(1 - np.exp(-3 * (t[0] + 1)) - v2) < 1e-13
        W = 4 * vt[4, s] / (1 - v2)
        assert abs(om.w0(u[:, s]) * np.exp(t[1] / 25.0) / W - 1) < 1e-13
        assert abs(t[4] - (2 * v2 - 1)) < 1e-15


@pytest.mark.parametrize("coords,t", [(1, 1.0), (2, 1.7)])
def test_refine_reaches_the_reference_root(oracle, om, coords, t):
    """From a W guess 1e-4 off, two Newton iterations reach the
    primitives the state was built from to 1e-11 relative, on a Minkowski and
    a Bjorken metric; the reference's algorithm 5 reaches the same root only
    to its own tolerance (v^2 to about 1e-12 absolute, so p to ~1e-12 gamma^2
    relative, up to ~1e-9 at gamma = 20)."""
    _mhd_oracle(oracle, coords, t)
    u, vt, vw = _states(200, seed=11)
    g33 = t * t if coords == 2 else 1.0
    gp = t if coords == 2 else 1.0
    rng = np.random.default_rng(1)
    nacc = 0
    for s in range(200):
        vp = vt[:, s].copy()
        if coords == 2:
            vp[3] /= t          # contravariant v^eta
            vp[7] /= t          # contravariant B^eta
        e, us, _ = oracle.prim2cons(vp)
        assert e == 0
        tg = om.targets(us / gp, vp, g33)
        y = (tg + 25 * 1e-4 * rng.standard_normal(5)).astype(np.float32)   # W off by ~1e-4
        st, vr = om.refine(us, y, g33=g33, gm=1.0 / gp, iters=2, tol=1e-11)
        guess = vw[:, s].copy()
        if coords == 2:
            guess[3] /= t; guess[7] /= t
        e, vo, _ = oracle.cons2prim(us, guess)
        assert e == 0
        if st == 0:
            nacc += 1
            sc = _scales(vp)
            assert np.all(np.abs(vr - vp) <= 1e-11 * sc), (s, vr, vp)
            assert np.all(np.abs(vo - vp) <= 2e-9 * sc), (s, vo, vp)
    assert nacc >= 190


def test_refinement_residual(om):
    """At the exact solution the residual vanishes and x(W) = v^2; its
    derivative matches central differences."""
    u, vt, _ = _states(60, seed=15)
    for s in range(60):
        us = u[:, s]
        S2 = (us[1:4] ** 2).sum()
        B2 = (us[5:8] ** 2).sum()
        SB2 = (us[1:4] @ us[5:8]) ** 2
        v2 = (vt[1:4, s] ** 2).sum()
        W = 4 * vt[4, s] / (1 - v2)
        f, df, x = om.e3p_w(W, S2, B2, SB2, us[4])
        assert abs(f) <= 1e-12 * W ** 3 and abs(x - v2) <= 1e-13
        for Wt in (0.9 * W, 1.3 * W):
            h = 1e-6 * Wt
            fd = (om.e3p_w(Wt + h, S2, B2, SB2, us[4])[0] -
                  om.e3p_w(Wt - h, S2, B2, SB2, us[4])[0]) / (2 * h)
            d = om.e3p_w(Wt, S2, B2, SB2, us[4])[1]
            assert abs(fd - d) <= 1e-6 * abs(d), (s, fd, d)


def test_gate_rejects_unconverged_and_accepts_converged(om):
    u, vt, _ = _states(100, seed=12)
    rej = acc = 0
    for s in range(100):
        tg = om.targets(u[:, s], vt[:, s])
        y = (tg + np.array([0, 25 * 3e-3, 0, 0, 0])).astype(np.float32)   # W off by 3e-3
        st0, _ = om.refine(u[:, s], y, iters=0, tol=1e-11)
        st3, _ = om.refine(u[:, s], y, iters=4, tol=1e-11)
        rej += st0 == 1
        acc += st3 == 0
    assert rej == 100 and acc >= 98


def test_features_invariances(om):
    u, _, _ = _states(40, seed=13)
    rng = np.random.default_rng(4)
    for s in range(40):
        x0 = om.features(u[:, s])
        lam = np.exp(rng.uniform(-3, 3))
        us = u[:, s].copy()
        us[:5] *= lam
        us[5:8] *= np.sqrt(lam)
        assert np.allclose(om.features(us), x0, rtol=0, atol=2e-6)
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        ur = u[:, s].copy()
        ur[1:4] = q @ ur[1:4]
        ur[5:8] = q @ ur[5:8]
        assert np.allclose(om.features(ur), x0, rtol=0, atol=2e-6)


def test_generator_prim2cons_matches_oracle(oracle):
    _mhd_oracle(oracle)
    u, vt, _ = _states(100, seed=14)
    for s in range(100):
        e, uo, _ = oracle.prim2cons(vt[:, s].copy())
        assert e == 0
        assert np.allclose(u[:, s], uo, rtol=1e-14, atol=1e-14 * np.abs(uo).max())


def test_generator_ranges():
    from echoamd.surrogate import mhd_e3p_states
    u, vt, vw = mhd_e3p_states(20000, _snap(), seed=2)
    gam = 1 / np.sqrt(1 - (vt[1:4] ** 2).sum(0))
    sig = (vt[5:8] ** 2).sum(0) / (4 * vt[4])
    assert gam.min() >= 1.0 and 18 < gam.max() < 22.5
    assert 0.7e-4 < sig.min() < 2e-4 and 0.7 < sig.max() < 1.5
    assert np.all(np.isfinite(u))


def test_trained_network_gain_cpu(om):
    """The committed network (run through the oracle's half-precision forward pass)
    against the method without it: the share of states whose two Newton
    iterations pass the 1e-11 gate, from the network's W versus from the
    state at rest (W = 4p, p = (U - B^2/2)/3)."""
    import os
    from echoamd.surrogate import WEIGHTS, load_weights
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights")
    P = load_weights()
    u, vt, _ = _states(2000, seed=16)
    B2 = (u[5:8] ** 2).sum(0)
    Wc = 4.0 * np.maximum(u[4] - 0.5 * B2, 0.1 * u[4]) / 3.0
    y1c = [25.0 * np.log(Wc[s] / om.w0(u[:, s])) for s in range(u.shape[1])]
    acc_net = acc_cold = 0
    worst = 0.0
    for s in range(u.shape[1]):
        y = om.forward(P, om.features(u[:, s]), 1)
        st, v = om.refine(u[:, s], y, iters=2, tol=1e-11)
        if st == 0:
            acc_net += 1
            worst = max(worst, float(np.max(np.abs(v - vt[:, s]) / _scales(vt[:, s]))))
        yc = np.array([0.0, y1c[s], 0, 0, 0], dtype=np.float32)
        acc_cold += om.refine(u[:, s], yc, iters=2, tol=1e-11)[0] == 0
    print(f"\naccepted after 2 iterations: network {acc_net}/2000, at-rest guess {acc_cold}/2000; "
          f"max error of accepted states {worst:.1e}")
    assert worst <= 1e-10
    assert acc_net >= 0.95 * 2000
    assert acc_net - acc_cold >= 0.3 * 2000
