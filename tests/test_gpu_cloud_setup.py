"""GPU cloud set-up: stencil selection (mp_knn_stencils) against the host kNN
and the oracle's brute force; reverse Cuthill-McKee order (mp_rcm_order)
against the host order and the oracle's restatement.

Both kNN sides return the k smallest (squared distance, index) pairs with the
same squared-distance arithmetic (no contraction), so the CSR must agree entry
for entry, ties included (an unjittered lattice has many exact ties).  The
RCM order is a permutation fixed by the stencil graph alone, so it must agree
exactly too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _host_knn(xyz, bnd, ks):
    import memphys_amd as M
    nv, dim = xyz.shape
    rp = np.zeros(nv + 1, dtype=np.int32)
    col = np.zeros(int(ks.sum()), dtype=np.int32)
    rc = M.host_lib().mph_knn(nv, dim, np.ascontiguousarray(xyz).ravel(), bnd,
                              np.ascontiguousarray(ks, dtype=np.int32), rp, col, 0)
    assert rc == 0
    return rp, col


def _compare(ctx, name, xyz, bnd, ks, brute=False):
    rp_g, col_g = ctx.knn_stencils(xyz, bnd, ks)
    rp_h, col_h = _host_knn(xyz, bnd, ks)
    same = np.array_equal(rp_g, rp_h) and np.array_equal(col_g, col_h)
    rows_off = 0 if same else int(sum(
        not np.array_equal(col_g[rp_h[i]:rp_h[i + 1]], col_h[rp_h[i]:rp_h[i + 1]])
        for i in range(len(ks))))
    _print(f"{name}: nv {len(ks)} entries {len(col_h)} identical {same} rows differing {rows_off}")
    assert same
    assert col_g.min() >= 0
    if brute:
        import oracle as O
        rp_b, col_b = O.knn_bruteforce(xyz, bnd, ks)
        assert np.array_equal(rp_g, rp_b) and np.array_equal(col_g, col_b)


@pytest.mark.parametrize("dim,n,jitter,k,layers", [
    (2, 40, 0.25, 30, (None, None, None)),
    (3, 16, 0.3, 40, (60, 55, 50)),
    (2, 30, 0.0, 30, (None, None, None)),   # lattice: exact distance ties
    (3, 12, 0.0, 40, (60, 55, 50)),
    (3, 8, 0.25, 112, (None, None, None)),  # stencils above 96: the larger candidate lists
    (2, 20, 0.0, 108, (None, None, None)),
])
def test_knn_matches_host(dim, n, jitter, k, layers):
    import memphys_amd as M
    c = M.box_cloud(dim=dim, n=n, jitter=jitter, seed=5, k_interior=k, k_layers=layers, rcm=False)
    with M.Context(0) as ctx:
        _compare(ctx, f"box dim {dim} n {n} jitter {jitter}", c.xyz, c.boundary, c.ksize,
                 brute=(c.nv <= 2000))


def test_knn_gmsh_annulus_and_duplicates():
    import os
    import memphys_amd as M
    path = os.path.join(os.path.dirname(__file__), "golden", "ecc_ratio_0.5_r_i_0.6_n_200.msh")
    c = M.gmsh_cloud(path, 20, rcm=False)
    with M.Context(0) as ctx:
        _compare(ctx, "gmsh annulus k 20", c.xyz, c.boundary, c.ksize, brute=True)
        # exact duplicates of some interior nodes: the node itself still leads
        # its stencil, its twin of lower index follows
        rng = np.random.default_rng(3)
        inner = np.flatnonzero(c.boundary == 0)
        dup = rng.choice(inner, size=25, replace=False)
        xyz = np.concatenate([c.xyz, c.xyz[dup]])
        bnd = np.concatenate([c.boundary, np.zeros(len(dup), dtype=np.uint8)])
        ks = np.full(len(bnd), 20, dtype=np.int32)
        _compare(ctx, "gmsh annulus with 25 duplicated nodes", xyz, bnd, ks)
        rp, col = ctx.knn_stencils(xyz, bnd, ks)
        assert all(col[rp[i]] == i for i in range(len(bnd)))


def test_knn_full_size_against_host():
    """The 1M-node cloud of the time-stepping workload (stencils 40/50/55/60)."""
    import time
    import memphys_amd as M
    c = M.box_cloud(dim=3, n=100, jitter=0.3, seed=7, k_interior=40, k_layers=(60, 55, 50),
                    rcm=False)
    with M.Context(0) as ctx:
        w = M.box_cloud(dim=3, n=10, jitter=0.3, seed=2, k_interior=40, rcm=False)
        ctx.knn_stencils(w.xyz, w.boundary, w.ksize)  # warm-up (module load)
        t0 = time.perf_counter()
        rp_g, col_g = ctx.knn_stencils(c.xyz, c.boundary, c.ksize)
        t1 = time.perf_counter()
    _print(f"1M-node kNN on the GPU: {t1 - t0:.3f}s incl. transfers")
    assert np.array_equal(rp_g, c.row_ptr) and np.array_equal(col_g, c.col)


def test_knn_rejects_bad_sizes():
    import memphys_amd as M
    c = M.box_cloud(dim=2, n=10, jitter=0.2, seed=1, k_interior=12, rcm=False)
    with M.Context(0) as ctx:
        with pytest.raises(M.MemphysError):
            ctx.knn_stencils(c.xyz, c.boundary, np.full(c.nv, 513, dtype=np.int32))
        with pytest.raises(M.MemphysError):
            ctx.knn_stencils(c.xyz, c.boundary, np.full(c.nv, c.nv + 1, dtype=np.int32))
        with pytest.raises(M.MemphysError):
            ctx.knn_stencils(c.xyz, c.boundary, np.zeros(c.nv, dtype=np.int32))


def _host_rcm(rp, col):
    import memphys_amd as M
    order = np.zeros(len(rp) - 1, dtype=np.int32)
    assert M.host_lib().mph_rcm(len(rp) - 1, rp, col, order) == 0
    return order


@pytest.mark.parametrize("dim,n,k", [(2, 40, 30), (3, 16, 40), (2, 30, 12)])
def test_rcm_matches_host(dim, n, k):
    import memphys_amd as M
    import oracle as O
    c = M.box_cloud(dim=dim, n=n, jitter=0.3, seed=9, k_interior=k, rcm=False)
    nv = c.nv
    # the graph itself, and two disjoint copies (a restart at the second)
    rp2 = np.concatenate([c.row_ptr, c.row_ptr[1:] + c.row_ptr[-1]]).astype(np.int32)
    col2 = np.concatenate([c.col, c.col + nv]).astype(np.int32)
    with M.Context(0) as ctx:
        for name, rp, col in (("graph", c.row_ptr, c.col), ("two components", rp2, col2)):
            g = ctx.rcm_order(rp, col)
            h = _host_rcm(rp, col)
            _print(f"RCM dim {dim} n {n} {name}: nv {len(rp) - 1} identical {np.array_equal(g, h)}")
            assert np.array_equal(g, h)
            assert np.array_equal(np.sort(g), np.arange(len(rp) - 1))
        if nv <= 2000:
            assert np.array_equal(ctx.rcm_order(c.row_ptr, c.col),
                                  O.rcm_reference(c.row_ptr, c.col))


def test_rcm_gmsh_and_full_size():
    import os
    import time
    import memphys_amd as M
    path = os.path.join(os.path.dirname(__file__), "golden", "ecc_ratio_0.5_r_i_0.6_n_200.msh")
    a = M.gmsh_cloud(path, 20, rcm=False)
    c = M.box_cloud(dim=3, n=100, jitter=0.3, seed=7, k_interior=40, k_layers=(60, 55, 50),
                    rcm=False)
    with M.Context(0) as ctx:
        assert np.array_equal(ctx.rcm_order(a.row_ptr, a.col), _host_rcm(a.row_ptr, a.col))
        t0 = time.perf_counter()
        g = ctx.rcm_order(c.row_ptr, c.col)
        t1 = time.perf_counter()
    h = _host_rcm(c.row_ptr, c.col)
    t2 = time.perf_counter()
    _print(f"1M-node RCM: GPU {t1 - t0:.3f}s incl. transfers, host {t2 - t1:.3f}s, "
           f"identical {np.array_equal(g, h)}")
    assert np.array_equal(g, h)


def test_rcm_rejects_bad_graph():
    import memphys_amd as M
    rp = np.array([0, 2, 3], dtype=np.int32)
    with M.Context(0) as ctx:
        with pytest.raises(M.MemphysError):
            ctx.rcm_order(rp, np.array([0, 5, 1], dtype=np.int32))
        with pytest.raises(M.MemphysError):
            ctx.rcm_order(rp, np.array([0, -1, 1], dtype=np.int32))
