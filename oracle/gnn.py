"""ORACLE -- test infrastructure only: the learned-guess network in float64.

Restates the network the HIP kernels evaluate (memphys_amd/csrc/gnn.hpp) with
numpy, so that the GPU forward pass, loss and parameter gradient can be
checked.  Weights are rounded to BF16 exactly as the device does (round to
nearest even) and then used in float64; features are rounded to float32 as the
device stores them; the hidden states a layer aggregates over the stencils are
rounded to BF16 (the device gathers a BF16 copy of them), and the gradient
treats that rounding as the identity, as the device's backward pass does.  The backward pass here is itself pinned by finite
differences (tests/test_oracle_gnn.py).
"""
from __future__ import annotations

import numpy as np

H, F, L = 64, 5, 3
W0 = 0
B0 = W0 + F * H
LAYER = B0 + H
LAYER_SIZE = 2 * H * H + H
WO = LAYER + L * LAYER_SIZE
BO = WO + H
NPARAM = BO + 1


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(theta: np.ndarray) -> dict:
    t = np.asarray(theta, dtype=np.float64)
    p = {"W0": t[W0:W0 + F * H].reshape(F, H), "b0": t[B0:B0 + H],
         "wo": t[WO:WO + H], "bo": t[BO]}
    for l in range(L):
        o = LAYER + l * LAYER_SIZE
        p[f"Ws{l}"] = t[o:o + H * H].reshape(H, H)
        p[f"Wn{l}"] = t[o + H * H:o + 2 * H * H].reshape(H, H)
        p[f"bl{l}"] = t[o + 2 * H * H:o + 2 * H * H + H]
    return p


def operators(A_bordered, nv: int):
    """(A~ node block as float32 values held in float64 CSR, D^-1) from the
    FP64 pressure operator (caller order, bordered or not)."""
    import scipy.sparse as sp
    A = A_bordered.tocsr()
    d = A.diagonal()[:nv]
    dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 1.0)
    An = A[:nv, :nv].tocsr()
    scaled = An.multiply(dinv[None, :]).tocsr()
    scaled.data = scaled.data.astype(np.float32).astype(np.float64)
    return sp.csr_matrix(scaled), dinv


def features(A_bordered, nv, p1, p2, b, xyz, dinv):
    n = A_bordered.shape[0]
    pext = 2.0 * p1 - p2
    if n > nv:
        pext = pext.copy()
        pext[nv:] = p1[nv:]
    r = b - A_bordered @ pext
    dp = p1[:nv] - p2[:nv]
    s = np.sqrt(np.mean(dp * dp))
    s = s if s > 1e-300 else 1.0
    X = np.zeros((nv, F))
    X[:, 0] = dp / s
    X[:, 1] = dinv * r[:nv] / s
    X[:, 2:2 + xyz.shape[1]] = xyz
    return X.astype(np.float32).astype(np.float64), s, pext[:nv]


def forward(theta, X, At, round_h=True):
    """round_h=False: the unrounded network, whose exact gradient the backward
    formulas are (the finite-difference check)."""
    p = unpack(theta)
    zs, hs, ms = [], [], []
    z = X @ p["W0"] + p["b0"]
    h = np.maximum(z, 0.0)
    zs.append(z)
    hs.append(h)
    ms.append(None)
    for l in range(L):
        m = At @ (bf16(h) if round_h else h)
        z = h @ p[f"Ws{l}"] + m @ p[f"Wn{l}"] + p[f"bl{l}"]
        h = np.maximum(z, 0.0)
        zs.append(z)
        hs.append(h)
        ms.append(m)
    d = h @ p["wo"] + p["bo"]
    return d, (zs, hs, ms)


def loss_and_grad(theta, X, At, target, round_h=True):
    p = unpack(theta)
    d, (zs, hs, ms) = forward(theta, X, At, round_h)
    nv = X.shape[0]
    e = d - target
    loss = np.mean(e * e)
    dd = 2.0 * e / nv
    g = np.zeros(NPARAM)
    g[WO:WO + H] = hs[L].T @ dd
    g[BO] = dd.sum()
    dh = np.outer(dd, p["wo"])
    for l in range(L, 0, -1):
        dz = dh * (zs[l] > 0.0)
        o = LAYER + (l - 1) * LAYER_SIZE
        g[o:o + H * H] = (hs[l - 1].T @ dz).ravel()
        g[o + H * H:o + 2 * H * H] = (ms[l].T @ dz).ravel()
        g[o + 2 * H * H:o + 2 * H * H + H] = dz.sum(axis=0)
        dh = dz @ p[f"Ws{l - 1}"].T + At.T @ (dz @ p[f"Wn{l - 1}"].T)
    dz = dh * (zs[0] > 0.0)
    g[W0:W0 + F * H] = (X.T @ dz).ravel()
    g[B0:B0 + H] = dz.sum(axis=0)
    return loss, g, d


def evaluate(theta_f32, A_bordered, nv, p1, p2, b, xyz, p_true=None):
    """Guess (nv), and with p_true the loss and gradient, as mp_guess_eval."""
    theta = bf16(theta_f32)
    At, dinv = operators(A_bordered, nv)
    X, s, pext = features(A_bordered, nv, p1, p2, b, xyz, dinv)
    if p_true is None:
        d, _ = forward(theta, X, At)
        return pext + s * d, None, None
    target = (p_true[:nv] - pext) / s
    loss, g, d = loss_and_grad(theta, X, At, target)
    return pext + s * d, g, loss
