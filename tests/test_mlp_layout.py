"""The parameter layout of the learned con2prim as include/echo_amd.h
states it (ECHO_MLP_NPARAM floats: W1[64][8] b1[64] W2[64][64] b2[64]
W3[64][64] b3[64] W4[5][64] b4[5], row-major [out][in]), checked with an
independent numpy forward pass (plain matrix products, hard-swish
h = z clamp(z / 6 + 1/2, 0, 1)) against the oracle's FP64 forward pass, on
random parameters and on the shipped weights.  The GPU's forward pass is
compared with the oracle's in tests/test_gpu_c2p_mlp.py, so this closes the
chain header -> numpy -> oracle -> GPU.  A transposed reading of the
weights, or a shifted offset, fails."""
import os

import numpy as np
import pytest

from libs import PKG, OracleMLP

SHAPES = [(64, 8), (64,), (64, 64), (64,), (64, 64), (64,), (5, 64), (5,)]


def unpack(P):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(P[o:o + n].astype(np.float64).reshape(s))
        o += n
    assert o == len(P) == 9221
    return out


def hswish(z):
    return z * np.clip(z * np.float64(np.float32(1 / 6)) + 0.5, 0.0, 1.0)


def numpy_forward(P, X, transpose=False):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(P)
    if transpose:   # a mistaken [in][out] reading of the square layers
        W2, W3 = W2.T, W3.T
    h = X.astype(np.float64)
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        h = hswish(h @ W.T + b)
    return h @ W4.T + b4


def _params():
    rng = np.random.default_rng(3)
    yield "random", (rng.standard_normal(9221) * 0.3).astype(np.float32)
    path = os.path.join(PKG, "weights", "c2p_e3p_mlp.npy")
    if os.path.exists(path):
        yield "trained", np.load(path).astype(np.float32)


@pytest.mark.parametrize("which", ["random", "trained"])
def test_forward_matches_header_layout(which):
    P = dict(_params()).get(which)
    if P is None:
        pytest.skip("no trained weights")
    om = OracleMLP()
    X = np.random.default_rng(5).uniform(-2, 2, (200, 8)).astype(np.float32)
    ref = np.array([om.forward(P, x, rnd=0) for x in X], dtype=np.float64)
    mine = numpy_forward(P, X)
    scale = np.abs(ref).max()
    assert np.abs(mine - ref).max() <= 1e-6 * scale
    # the check sees a transposed reading
    assert np.abs(numpy_forward(P, X, transpose=True) - ref).max() > 1e-3 * scale
