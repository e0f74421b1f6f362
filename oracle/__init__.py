"""ORACLE -- test infrastructure only.

ctypes access to ``oracle/_build/liboracle.so``, the FP64 CPU restatement of
MeMPhyS's fractional-step path (see oracle.cpp), plus small pure-Python
restatements (brute-force kNN, RCM) used to check the product's set-up code.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.

Parity pins and gaps are listed in oracle.cpp's header.  The kNN tie order of
nanoflann (nanoflann.hpp:175-201) is not restated: ties are broken by index,
as in the product; parity unpinned for exactly equidistant points.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.or_poly_exponents.argtypes = [C.c_int, C.c_int, _ip]
        L.or_poly_exponents.restype = C.c_int
        L.or_stencil_weights.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _ip, _ip, C.c_int,
                                         _vp, C.c_int, _dp, _dp, _vp, _dp, _vp, C.c_int, C.c_int]
        L.or_stencil_weights.restype = C.c_int
        L.or_single_stencil.argtypes = [C.c_int, C.c_int, C.c_int, _ip, C.c_int, _dp, _dp, _dp,
                                        _vp, _dp, C.c_int]
        L.or_single_stencil.restype = C.c_double
        L.or_spmv.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.or_spmv.restype = None
        L.or_fs_create.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _vp, _dp, _dp, _u8, _u8,
                                   _u8, _vp, _u8, C.c_double, C.c_double, C.c_double, C.c_int,
                                   C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double]
        L.or_fs_create.restype = _vp
        L.or_fs_destroy.argtypes = [_vp]
        L.or_fs_destroy.restype = None
        L.or_fs_system_size.argtypes = [_vp]
        L.or_fs_system_size.restype = C.c_int
        L.or_fs_ppe_nnz.argtypes = [_vp]
        L.or_fs_ppe_nnz.restype = C.c_int
        L.or_fs_ppe_export.argtypes = [_vp, _ip, _ip, _dp]
        L.or_fs_ppe_export.restype = None
        L.or_ppe_solve.argtypes = [_vp, _dp, _dp, _dp, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        L.or_ppe_solve.restype = C.c_int
        L.or_fs_set_body_force.argtypes = [_vp, _vp, _vp, _vp]
        L.or_fs_set_body_force.restype = None
        L.or_fs_step.argtypes = [_vp, C.c_int, _dp, _dp, _vp, _dp]
        L.or_fs_step.restype = C.c_int
        L.or_fs_predict.argtypes = [_vp, C.c_int, _dp, _dp, _vp, _dp, _dp, _dp]
        L.or_fs_predict.restype = None
        L.or_fs_correct.argtypes = [_vp, _dp, _dp, _vp, _dp, _dp, _dp]
        L.or_fs_correct.restype = None
        L.or_fs_last.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.or_fs_last.restype = None
        L.or_divergence.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _vp, _u8, _dp, _dp, _vp,
                                    _vp, _dp]
        L.or_divergence.restype = None
        L.or_set_threads.argtypes = [C.c_int]
        L.or_coeff_export.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _vp, _dp, _dp, _u8,
                                      _u8, C.c_double, C.c_double, C.c_double, _vp, _vp, _vp]
        L.or_coeff_export.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def poly_exponents(dim: int, poly_deg: int) -> np.ndarray:
    m = (poly_deg + 1) * (poly_deg + 2) // 2 if dim == 2 else \
        (poly_deg + 1) * (poly_deg + 2) * (poly_deg + 3) // 6
    out = np.zeros(m * dim, dtype=np.int32)
    got = lib().or_poly_exponents(dim, poly_deg, out)
    assert got == m
    return out.reshape(m, dim)


def single_stencil(X: np.ndarray, phs_deg: int, poly_deg: int, refine: int = 0):
    """Weights of one stencil (rows of X, center first): (gx, gy, gz|None, lap, cond)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    k, dim = X.shape
    ex = np.ascontiguousarray(poly_exponents(dim, poly_deg).ravel())
    gx, gy, lap = np.zeros(k), np.zeros(k), np.zeros(k)
    gz = np.zeros(k) if dim == 3 else None
    cond = lib().or_single_stencil(dim, phs_deg, len(ex) // dim, ex, k, X.ravel(), gx, gy,
                                   _ptr(gz), lap, refine)
    return gx, gy, gz, lap, cond


def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP loops; results do not depend on it."""
    lib().or_set_threads(int(n))


def stencil_weights(cloud, phs_deg: int, poly_deg: int, rows=None, refine: int = 0,
                    threads: int = 0, want_cond: bool = False):
    """CSR-aligned weights (gx, gy, gz|None, lap) for all rows or the listed rows."""
    dim = cloud.dim
    ex = np.ascontiguousarray(poly_exponents(dim, poly_deg).ravel())
    W = cloud.nnz
    gx, gy, lap = np.zeros(W), np.zeros(W), np.zeros(W)
    gz = np.zeros(W) if dim == 3 else None
    r = None
    n = cloud.nv
    if rows is not None:
        r = np.ascontiguousarray(rows, dtype=np.int32)
        n = len(r)
    cond = np.zeros(n) if want_cond else None
    rc = lib().or_stencil_weights(dim, phs_deg, len(ex) // dim, ex,
                                  np.ascontiguousarray(cloud.xyz.ravel()), cloud.row_ptr,
                                  cloud.col, n, _ptr(r), cloud.nv, gx, gy, _ptr(gz), lap,
                                  _ptr(cond), refine, threads)
    assert rc == 0
    return gx, gy, gz, lap, cond


def spmv(cloud, val, x):
    y = np.zeros(cloud.nv)
    lib().or_spmv(cloud.nv, cloud.row_ptr, cloud.col, np.ascontiguousarray(val),
                  np.ascontiguousarray(x, dtype=np.float64), y)
    return y


class FractionalStep:
    """FRACTIONAL_STEP_1 restated on the CPU (solver_kind 0 BiCGSTAB, 1 dense LU;
    precond 0 none, 1 Jacobi, 2 ILUT)."""

    def __init__(self, cloud, w, u_dir, v_dir, w_dir, p_dir, rho, mu, dt, order,
                 solver_kind=0, precond=2, tol=1e-12, maxit=5000, droptol=1e-4, fill=10.0):
        gx, gy, gz, lap = w
        self.cloud = cloud
        self._keep = [np.ascontiguousarray(a) if a is not None else None
                      for a in (gx, gy, gz, lap, u_dir, v_dir, w_dir, p_dir)]
        gx, gy, gz, lap, u_dir, v_dir, w_dir, p_dir = self._keep
        self.h = lib().or_fs_create(
            cloud.nv, cloud.dim, cloud.row_ptr, cloud.col, gx, gy, _ptr(gz), lap,
            np.ascontiguousarray(cloud.normal.ravel()), cloud.boundary,
            u_dir.astype(np.uint8), v_dir.astype(np.uint8),
            _ptr(w_dir.astype(np.uint8) if w_dir is not None else None), p_dir.astype(np.uint8),
            rho, mu, dt, order, solver_kind, precond, tol, maxit, droptol, fill)
        self.n_sys = lib().or_fs_system_size(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_fs_destroy(self.h)
            self.h = None

    def ppe_csr(self):
        nnz = lib().or_fs_ppe_nnz(self.h)
        rp = np.zeros(self.n_sys + 1, dtype=np.int32)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        lib().or_fs_ppe_export(self.h, rp, ci, v)
        return rp, ci, v

    def solve(self, rhs, guess):
        xn = np.zeros(self.n_sys)
        xo = np.ascontiguousarray(guess, dtype=np.float64).copy()
        rr, ra = C.c_double(), C.c_double()
        it = lib().or_ppe_solve(self.h, np.ascontiguousarray(rhs, dtype=np.float64), xn, xo,
                                C.byref(rr), C.byref(ra))
        return xn, xo, it, rr.value

    def set_body_force(self, fx=None, fy=None, fz=None):
        """Body force added to the explicit source of the following steps
        (fractional_step_1.cpp:104-118); no arguments removes it."""
        if fx is None:
            lib().or_fs_set_body_force(self.h, None, None, None)
            return
        self._force = [np.ascontiguousarray(f, dtype=np.float64) if f is not None else None
                       for f in (fx, fy, fz)]
        lib().or_fs_set_body_force(self.h, *[_ptr(f) for f in self._force])

    def step(self, it, u, v, w, p):
        """Advances in place (u, v, w, p hold old values on entry)."""
        return lib().or_fs_step(self.h, it, u, v, _ptr(w), p)

    def predict(self, it, u, v, w, p):
        """Predictor and pressure right-hand side: (uh (dim, nv), b)."""
        nv, dim = self.cloud.nv, self.cloud.dim
        uh = np.zeros(dim * nv)
        b = np.zeros(self.n_sys)
        w_ = np.ascontiguousarray(w, dtype=np.float64) if w is not None else None
        lib().or_fs_predict(self.h, it, np.ascontiguousarray(u, dtype=np.float64),
                            np.ascontiguousarray(v, dtype=np.float64), _ptr(w_),
                            np.ascontiguousarray(p, dtype=np.float64), uh, b)
        return uh.reshape(dim, nv), b

    def correct(self, u, v, w, uh, p):
        """Corrector with the boundary fix-up: u_new (dim, nv)."""
        nv, dim = self.cloud.nv, self.cloud.dim
        un = np.zeros(dim * nv)
        w_ = np.ascontiguousarray(w, dtype=np.float64) if w is not None else None
        lib().or_fs_correct(self.h, np.ascontiguousarray(u, dtype=np.float64),
                            np.ascontiguousarray(v, dtype=np.float64), _ptr(w_),
                            np.ascontiguousarray(uh, dtype=np.float64).ravel(),
                            np.ascontiguousarray(p, dtype=np.float64), un)
        return un.reshape(dim, nv)

    def last(self):
        it, rr, ra = C.c_int(), C.c_double(), C.c_double()
        lib().or_fs_last(self.h, C.byref(it), C.byref(rr), C.byref(ra))
        return it.value, rr.value, ra.value


def coeff_matrix(cloud, w, dirichlet, unsteady=0.0, conv=0.0, diff=1.0):
    """SOLVER::init + calc_coeff_matrix (solver.cpp:29-38, 165-214) as a scipy
    CSR matrix: n = nv + 1 (full Neumann, bordered) or nv."""
    import scipy.sparse as sp
    gx, gy, gz, lap = (np.ascontiguousarray(a) if a is not None else None for a in w)
    d = np.ascontiguousarray(dirichlet, dtype=np.uint8)
    args = (cloud.nv, cloud.dim, cloud.row_ptr, cloud.col, gx, gy, _ptr(gz), lap,
            np.ascontiguousarray(cloud.normal.ravel()), cloud.boundary, d, float(unsteady),
            float(conv), float(diff))
    nnz = lib().or_coeff_export(*args, None, None, None)
    n_max = cloud.nv + 1
    rp = np.zeros(n_max + 1, dtype=np.int32)
    ci = np.zeros(nnz, dtype=np.int32)
    v = np.zeros(nnz)
    n = lib().or_coeff_export(*args, _ptr(rp), _ptr(ci), _ptr(v))
    return sp.csr_matrix((v, ci, rp[:n + 1]), shape=(n, n))


def divergence(cloud, gx, gy, gz, u, v, w):
    out = np.zeros(6)
    res = np.zeros(cloud.nv)
    lib().or_divergence(cloud.nv, cloud.dim, cloud.row_ptr, cloud.col, gx, gy, _ptr(gz),
                        cloud.boundary, np.ascontiguousarray(u), np.ascontiguousarray(v),
                        _ptr(np.ascontiguousarray(w) if w is not None else None), _ptr(res), out)
    return out, res


# ----------------------------------------------------------------------------- set-up checks
def knn_bruteforce(xyz: np.ndarray, boundary: np.ndarray, ksize: np.ndarray):
    """CLOUD::calc_cloud_points_fast (cloud.cpp:570-648) by brute force: interior
    nodes take their k nearest of all nodes, boundary nodes themselves then
    their k-1 nearest interior nodes; order (squared distance, index)."""
    nv = xyz.shape[0]
    interior = np.flatnonzero(boundary == 0)
    rows = []
    for i in range(nv):
        d = xyz - xyz[i]
        d2 = np.zeros(nv)
        for a in range(xyz.shape[1]):
            d2 = d2 + d[:, a] * d[:, a]
        if boundary[i]:
            cand = interior
            order = np.lexsort((cand, d2[cand]))[: ksize[i] - 1]
            rows.append(np.concatenate([[i], cand[order]]))
        else:
            order = np.lexsort((np.arange(nv), d2))[: ksize[i]]
            rows.append(order)
    rp = np.zeros(nv + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return rp, np.concatenate(rows).astype(np.int32)


def rcm_reference(row_ptr: np.ndarray, col: np.ndarray) -> np.ndarray:
    """cuthill_mckee_ordering + reverse_cuthill_mckee_ordering
    (general_functions.cpp:807-879), restated for connected stencil graphs:
    returns order[old] = new."""
    nv = len(row_ptr) - 1
    deg = np.diff(row_ptr)
    start = int(np.argmin(deg))  # first index of minimum degree
    numbering = [start]
    visited = np.zeros(nv, dtype=bool)
    visited[start] = True
    index = 0
    while len(numbering) < nv:
        v = numbering[index]
        new_set = []
        for c in col[row_ptr[v]:row_ptr[v + 1]]:
            if not visited[c]:
                visited[c] = True
                new_set.append((int(deg[c]), int(c)))
        new_set.sort()
        numbering.extend(c for _, c in new_set)
        index += 1
    order = np.empty(nv, dtype=np.int64)
    order[np.array(numbering)] = np.arange(nv)
    return (nv - order - 1).astype(np.int32)
