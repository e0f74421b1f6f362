"""The float64 network oracle: its backward pass against finite differences,
and its BF16 rounding against the IEEE definition (CPU)."""
import numpy as np
import scipy.sparse as sp

from oracle import gnn as G


def _setup(seed=0, nv=40):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(nv), 6)
    cols = rng.integers(0, nv, size=rows.size)
    At = sp.csr_matrix((rng.standard_normal(rows.size) * 0.3, (rows, cols)), shape=(nv, nv))
    X = rng.standard_normal((nv, G.F))
    theta = rng.standard_normal(G.NPARAM) * 0.2
    target = rng.standard_normal(nv)
    return theta, X, At, target


def test_backward_matches_finite_differences():
    theta, X, At, target = _setup()
    loss, g, _ = G.loss_and_grad(theta, X, At, target)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(G.NPARAM, 60, replace=False),
                          [G.W0, G.B0, G.LAYER, G.LAYER + G.H * G.H, G.WO, G.BO]])
    eps = 1e-6
    for i in idx:
        tp = theta.copy()
        tp[i] += eps
        tm = theta.copy()
        tm[i] -= eps
        fd = (G.loss_and_grad(tp, X, At, target)[0] - G.loss_and_grad(tm, X, At, target)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_zero_output_layer_gives_extrapolation():
    theta, X, At, target = _setup(2)
    theta[G.WO:G.BO + 1] = 0.0
    d, _ = G.forward(theta, X, At)
    assert np.all(d == 0.0)


def test_bf16_rounding():
    # BF16 keeps 8 significand bits: the spacing above 1 is 2^-7
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -10, -3.14159, 1.0 + 3 * 2 ** -8],
                 dtype=np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                        # exact tie -> even (1.0)
    assert r[2] == 1.0 + 2 ** -7              # above the tie -> up
    assert r[4] == 1.0 + 2 ** -6              # tie between 1+2^-7 and 1+2^-6 -> even
    assert abs(r[3] - (-3.140625)) < 1e-12
