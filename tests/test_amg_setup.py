"""CPU checks of the multigrid set-up (host C++ builder) on the 3D cube's
pressure operator.

The builder's output is pinned by what the method fixes independently of
its code: the aggregates partition the nodes; the prolongator equals
(I - w D^-1 A) P_tent formed here with scipy from the aggregates; the coarse
operator equals P^T A P formed with scipy; the truncated prolongator and the
sparsified coarse operator equal their rules applied here to those, and keep
the constant mode (a truncation that skips the rescale does not); the aggregation follows the rule
it states (restated below on the same strength graph); the coarsest-level
solve is the inverse of the ones-bordered matrix (numpy LAPACK inverse).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import lid_bc


@pytest.fixture(scope="module")
def ppe_block(cube3d_small):
    import oracle as O
    c = cube3d_small
    w = O.stencil_weights(c, 3, 3, threads=8)[:4]
    u_dir, p_dir, _ = lid_bc(c, 2)
    fs = O.FractionalStep(c, w, u_dir, u_dir, u_dir, p_dir, 1.0, 0.0025, 5e-4, 2)
    rp, ci, v = fs.ppe_csr()
    A = sp.csr_matrix((v, ci, rp), shape=(fs.n_sys, fs.n_sys))[:c.nv, :c.nv].tocsr()
    A.sort_indices()
    return A


def _aggregate_restated(A, k):
    n = A.shape[0]
    rows, cols = [], []
    for i in range(n):
        idx = A.indices[A.indptr[i]:A.indptr[i + 1]]
        val = np.abs(A.data[A.indptr[i]:A.indptr[i + 1]])
        m = idx != i
        idx, val = idx[m], val[m]
        order = np.lexsort((idx, -val))[:k]  # largest |a|, ties to the smaller column
        rows += [i] * len(order)
        cols += list(idx[order])
    S = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    S = ((S + S.T) > 0).tocsr()
    S.sort_indices()
    agg = -np.ones(n, dtype=np.int64)
    na = 0
    for i in range(n):
        if agg[i] >= 0:
            continue
        nb = S.indices[S.indptr[i]:S.indptr[i + 1]]
        if np.all(agg[nb] < 0):
            agg[i] = na
            agg[nb] = na
            na += 1
    for i in range(n):
        if agg[i] >= 0:
            continue
        nb = S.indices[S.indptr[i]:S.indptr[i + 1]]
        a = agg[nb][agg[nb] >= 0]
        if len(a):
            agg[i] = a[0]
        else:
            agg[i] = na
            na += 1
    return agg, na


def test_aggregates_partition_and_follow_rule(ppe_block):
    import memphys_amd as M
    A = ppe_block
    agg, P, Ac = M.amg_first_level(A, k=6)
    n, na = A.shape[0], Ac.shape[0]
    assert agg.min() == 0 and agg.max() == na - 1
    assert np.all(np.bincount(agg, minlength=na) > 0)
    assert na < n / 4, "coarsening ratio"
    ref, ref_na = _aggregate_restated(A, 6)
    assert ref_na == na
    assert np.array_equal(ref, agg)


def test_prolongator_and_galerkin_operator(ppe_block):
    import memphys_amd as M
    A = ppe_block
    omega = 0.66
    agg, P, Ac = M.amg_first_level(A, k=6, omega_p=omega)
    n, na = A.shape[0], Ac.shape[0]
    Pt = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, na))
    P_ref = (Pt - omega * sp.diags(1.0 / A.diagonal()) @ (A @ Pt)).tocsr()
    dP = abs(P - P_ref).max()
    assert dP <= 1e-13 * abs(P_ref).max()
    # a planted transposition (A^T in place of A) must fail the same check
    P_bad = (Pt - omega * sp.diags(1.0 / A.diagonal()) @ (A.T @ Pt)).tocsr()
    assert abs(P - P_bad).max() > 1e-6 * abs(P_ref).max()
    Ac_ref = (P_ref.T @ A @ P_ref).tocsr()
    dA = abs(Ac - Ac_ref).max()
    assert dA <= 1e-12 * abs(Ac_ref).max()
    # the constant mode survives: row sums of the Neumann Laplacian are ~0 and
    # P reproduces constants, so the coarse operator annihilates constants too
    ones_c = np.ones(na)
    assert np.abs(P @ ones_c - 1.0).max() <= 1e-8
    assert np.abs(Ac @ ones_c).max() <= 1e-6 * abs(Ac).max()


def test_bordered_inverse(ppe_block):
    import memphys_amd as M
    A = ppe_block
    _, _, Ac = M.amg_first_level(A, k=6)
    _, _, A2 = M.amg_first_level(Ac, k=6)
    n = A2.shape[0]
    assert 10 <= n <= 700
    G = M.amg_bordered_inverse(A2)
    B = np.zeros((n + 1, n + 1))
    B[:n, :n] = A2.toarray()
    B[:n, n] = 1.0
    B[n, :n] = 1.0
    G_ref = np.linalg.inv(B)[:n, :n]
    assert np.abs(G - G_ref).max() <= 1e-9 * np.abs(G_ref).max()
    # x = G b solves A x + lambda 1 = b, sum(x) = 0: the residual is constant
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    x = G @ b
    assert abs(x.sum()) <= 1e-9 * np.abs(x).max()
    r = A2 @ x - b
    assert np.abs(r - r.mean()).max() <= 1e-8 * np.linalg.norm(b)


def test_dense_inverse(ppe_block):
    """The coarsest solve of a system with Dirichlet rows (nonsingular): the
    plain inverse, against LAPACK; a planted transpose fails."""
    import memphys_amd as M
    A = ppe_block
    _, _, Ac = M.amg_first_level(A, k=6)
    _, _, A2 = M.amg_first_level(Ac, k=6)
    n = A2.shape[0]
    D = A2.toarray() + np.diag(np.abs(A2.toarray()).sum(axis=1))  # nonsingular, non-symmetric
    import scipy.sparse as sp
    G = M.amg_dense_inverse(sp.csr_matrix(D))
    G_ref = np.linalg.inv(D)
    assert np.abs(G - G_ref).max() <= 1e-10 * np.abs(G_ref).max()
    assert np.abs(G @ D - np.eye(n)).max() <= 1e-10
    if np.abs(D - D.T).max() > 0:
        assert np.abs(G.T - G_ref).max() > 1e-6 * np.abs(G_ref).max()


def _truncate_restated(P0, trunc, rescale=True):
    """Each row: entries below trunc x its largest |p| dropped, the rest scaled
    to the full row's sum."""
    P0 = P0.tocsr()
    rows, cols, vals = [], [], []
    for i in range(P0.shape[0]):
        c = P0.indices[P0.indptr[i]:P0.indptr[i + 1]]
        v = P0.data[P0.indptr[i]:P0.indptr[i + 1]]
        keep = np.abs(v) >= trunc * np.abs(v).max()
        s = v[keep].sum()
        f = v.sum() / s if (rescale and abs(s) > 1e-300) else 1.0
        rows += [i] * int(keep.sum())
        cols += list(c[keep])
        vals += list(v[keep] * f)
    return sp.csr_matrix((vals, (rows, cols)), shape=P0.shape)


def _sparsify_restated(Ac, drop):
    """Off-diagonal entries below drop x sqrt(|a_ii a_jj|) moved onto the diagonal."""
    Ac = Ac.tocoo()
    d = np.abs(Ac.tocsr().diagonal())
    weak = (Ac.row != Ac.col) & (np.abs(Ac.data) < drop * np.sqrt(d[Ac.row] * d[Ac.col]))
    lump = np.bincount(Ac.row[weak], weights=Ac.data[weak], minlength=Ac.shape[0])
    S = sp.csr_matrix((Ac.data[~weak], (Ac.row[~weak], Ac.col[~weak])), shape=Ac.shape)
    return (S + sp.diags(lump)).tocsr()


def test_truncated_prolongator_and_sparsified_operator(ppe_block):
    import memphys_amd as M
    A = ppe_block
    trunc, drop = 0.15, 0.01
    _, P0, _ = M.amg_first_level(A, k=6)
    agg, P, Ac = M.amg_first_level(A, k=6, trunc=trunc, drop=drop)
    n, na = P.shape
    # P: the untruncated prolongator, truncated and rescaled row by row
    P_ref = _truncate_restated(P0, trunc)
    assert P.nnz == P_ref.nnz and P.nnz < 0.6 * P0.nnz
    assert abs(P - P_ref).max() <= 1e-13 * abs(P_ref).max()
    # row sums survive the truncation, so constants are still reproduced; a
    # truncation without the rescale does not reproduce them
    assert np.abs(P @ np.ones(na) - P0 @ np.ones(na)).max() <= 1e-12
    assert np.abs(P @ np.ones(na) - 1.0).max() <= 1e-8
    bad = _truncate_restated(P0, trunc, rescale=False)
    assert np.abs(bad @ np.ones(na) - 1.0).max() > 1e-3
    # R A P with the weak entries lumped onto the diagonal
    Ac_ref = _sparsify_restated((P_ref.T @ A @ P_ref).tocsr(), drop)
    Ac_ref.eliminate_zeros()
    assert abs(Ac - Ac_ref).max() <= 1e-12 * abs(Ac_ref).max()
    assert Ac.nnz < 0.5 * (P_ref.T @ A @ P_ref).nnz
    assert np.abs(Ac @ np.ones(na)).max() <= 1e-6 * abs(Ac).max()
