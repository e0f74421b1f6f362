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
    # the unrounded network: its loss is smooth in theta
    theta, X, At, target = _setup()
    loss, g, _ = G.loss_and_grad(theta, X, At, target, round_h=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(G.NPARAM, 60, replace=False),
                          [G.W0, G.B0, G.LAYER, G.LAYER + G.H * G.H, G.WO, G.BO]])
    eps = 1e-6
    for i in idx:
        tp = theta.copy()
        tp[i] += eps
        tm = theta.copy()
        tm[i] -= eps
        fd = (G.loss_and_grad(tp, X, At, target, round_h=False)[0] -
              G.loss_and_grad(tm, X, At, target, round_h=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_aggregated_states_are_rounded_to_bf16():
    """The rounded network differs from the unrounded one by the BF16 rounding
    of the aggregated states only: no difference when those are BF16 numbers
    already (a ReLU network on BF16-representable inputs with a zero first
    layer weight), a small one otherwise."""
    theta, X, At, target = _setup(3)
    d1, _ = G.forward(theta, X, At)
    d0, _ = G.forward(theta, X, At, round_h=False)
    rel = np.linalg.norm(d1 - d0) / np.linalg.norm(d0)
    assert 0 < rel < 2e-2
    th = theta.copy()
    th[G.W0:G.B0] = 0.0                       # h0 = relu(b0): the bias, BF16 already
    th[G.LAYER:G.LAYER + 2 * G.H * G.H] = 0.0  # layer 1: h1 = relu(bias), BF16 already
    th[G.LAYER + G.LAYER_SIZE:] = 0.0          # layers 2, 3 and the output: zero
    th = G.bf16(th.astype(np.float32))
    e1, (_, hs1, ms1) = G.forward(th, X, At)
    e0, (_, hs0, ms0) = G.forward(th, X, At, round_h=False)
    assert np.array_equal(ms1[1], ms0[1]) and np.array_equal(ms1[2], ms0[2])


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
