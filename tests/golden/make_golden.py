"""Writes tests/golden/stencil_weights.npz: golden stencil weights.

An independent numpy restatement of coefficient_computations.cpp:70-328
(A from pow(r^2, p/2) and pow-based monomials, right-hand sides from the
reference's formulas, LAPACK dgesv -- partial-pivot LU, the method of
Eigen's partialPivLu -- for the solve).  It shares no code with the oracle
(different language, pow instead of sqrt/products, LAPACK instead of its own
LU), so agreement pins the oracle's assembly and LU.

Run: python tests/golden/make_golden.py   (deterministic; seeds below)
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def exponents(dim, deg):
    # graded order of parameters.cpp:217-267, enumerated directly
    out = []
    for d in range(deg + 1):
        if dim == 2:
            for b in range(d + 1):
                out.append((d - b, b))
        else:
            for a in range(d, -1, -1):
                for b in range(d - a, -1, -1):
                    out.append((a, b, d - a - b))
    return np.array(out, dtype=np.int64)


def reference_weights(X, phs, deg):
    """Weights of stencil X (k x dim, center first) as the reference computes them."""
    X = np.array(X, dtype=np.float64)
    k, dim = X.shape
    E = exponents(dim, deg)
    m = len(E)
    mn, mx = X.min(axis=0), X.max(axis=0)
    sc = mx - mn
    xs = (X - mn) / sc
    M = k + m
    A = np.zeros((M, M))
    for i in range(k):
        for j in range(i + 1, k):
            r2 = np.sum((xs[i] - xs[j]) ** 2)
            A[i, j] = A[j, i] = np.power(r2, phs / 2.0)
    for i in range(k):
        for t in range(m):
            v = np.power(xs[i, 0], float(E[t, 0]))
            for a in range(1, dim):
                v = v * np.power(xs[i, a], float(E[t, a]))
            A[i, k + t] = A[k + t, i] = v
    R = np.zeros((M, dim + 1))
    c = xs[0]
    for i in range(k):
        r2 = np.sum((xs[i] - c) ** 2)
        for a in range(dim):
            R[i, a] = -phs * (xs[i, a] - c[a]) * np.power(r2, (phs - 2.0) / 2.0) / sc[a]
        if abs(r2) > 1e-14:
            s = 0.0
            for a in range(dim):
                d2 = (xs[i, a] - c[a]) ** 2
                g = phs * phs * d2 * np.power(r2, (phs - 4.0) / 2.0)
                g = g - 2.0 * phs * d2 * np.power(r2, (phs - 4.0) / 2.0)
                g = g + phs * np.power(r2, (phs - 2.0) / 2.0)
                s += g / (sc[a] * sc[a])
            R[i, dim] = s
    for t in range(m):
        e = E[t].astype(float)
        for a in range(dim):
            if e[a] > 0:
                v = e[a] * np.power(c[a], e[a] - 1.0)
                for b in range(dim):
                    if b != a:
                        v = v * np.power(c[b], e[b])
                R[k + t, a] = v / sc[a]
        s = 0.0
        for a in range(dim):
            if e[a] > 1:
                v = e[a] * (e[a] - 1.0) * np.power(c[a], e[a] - 2.0)
                for b in range(dim):
                    if b != a:
                        v = v * np.power(c[b], e[b])
                s += v / (sc[a] * sc[a])
        R[k + t, dim] = s
    W = np.linalg.solve(A, R)
    return W[:k].T.copy(), np.linalg.cond(A, 1)


def stencil(rng, dim, k, one_sided=False):
    """Center at 0.5 (or on a face) plus k-1 jittered lattice neighbours."""
    h = 0.05
    pts = []
    rad = 4
    grid = np.array(np.meshgrid(*[np.arange(-rad, rad + 1)] * dim, indexing="ij")).reshape(dim, -1).T
    if one_sided:
        grid = grid[grid[:, dim - 1] < 0]
    d = np.sum(grid.astype(float) ** 2, axis=1)
    grid = grid[np.argsort(d, kind="stable")]
    center = np.full(dim, 0.5)
    if one_sided:
        pts.append(center.copy())
        cand = grid[: k - 1]
    else:
        cand = grid[:k]
    for g in cand:
        if not one_sided and np.all(g == 0):
            pts.append(center.copy())
            continue
        pts.append(center + h * (g + rng.uniform(-0.3, 0.3, dim)))
    X = np.array(pts)
    if not one_sided:
        i0 = int(np.argmin(np.sum((X - center) ** 2, axis=1)))
        X[[0, i0]] = X[[i0, 0]]
    return X


def main():
    rng = np.random.default_rng(20240611)
    cases = [
        (2, 20, 3, 3, False), (2, 30, 3, 3, False), (2, 30, 3, 3, True), (2, 30, 5, 3, False),
        (2, 16, 3, 2, False), (3, 40, 3, 3, False), (3, 60, 3, 3, True), (3, 50, 3, 3, False),
        (3, 40, 5, 3, False), (3, 20, 7, 2, False),
    ]
    out = {}
    for idx, (dim, k, phs, deg, one_sided) in enumerate(cases):
        X = stencil(rng, dim, k, one_sided)
        W, cond = reference_weights(X, phs, deg)
        out[f"c{idx}_X"] = X
        out[f"c{idx}_meta"] = np.array([dim, k, phs, deg], dtype=np.int64)
        out[f"c{idx}_W"] = W
        out[f"c{idx}_cond1"] = np.array(cond)
    out["ncases"] = np.array(len(cases))
    np.savez(os.path.join(HERE, "stencil_weights.npz"), **out)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
