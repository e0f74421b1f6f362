"""memphys_amd: MI355X (gfx950) meshless RBF-FD fractional-step path.

Python access to the two native libraries built in-tree:

* ``libmemphys_hip.so``  -- HIP kernels behind the C ABI of ``include/memphys_hip.h``;
* ``libmemphys_host.so`` -- host set-up (jittered box clouds, kNN stencils, RCM).

Nothing here computes on the CPU in place of the GPU: every numerical call
goes to the HIP library, and loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libmemphys_hip.so")
HOST_LIB_PATH = os.path.join(_HERE, "libmemphys_host.so")

MP_U, MP_V, MP_W, MP_P = 0, 1, 2, 3
MP_SU, MP_SV, MP_SW = 4, 5, 6  # AB2 source history (exact restart)
MP_GX, MP_GY, MP_GZ, MP_LAP = 0, 1, 2, 3
MP_PREC_FP64, MP_PREC_MIXED = 0, 1
MP_GUESS_PREVIOUS, MP_GUESS_LEARNED, MP_GUESS_EXTRAPOLATE, MP_GUESS_PROJECT = 0, 1, 2, 3
MP_PRECOND_JACOBI, MP_PRECOND_AMG = 0, 1
MP_KRYLOV_BICGSTAB, MP_KRYLOV_GMRES = 0, 1

# Every symbol include/memphys_hip.h declares.
HIP_SYMBOLS = (
    "mp_ctx_create", "mp_ctx_destroy", "mp_last_error", "mp_sync", "mp_cloud_upload",
    "mp_weights_assemble", "mp_weights_set", "mp_operator_apply", "mp_solver_init",
    "mp_solver_system_size", "mp_solver_solve", "mp_solver_krylov", "mp_fs_create", "mp_fields_set",
    "mp_fields_get", "mp_fs_step", "mp_stats_get", "mp_guess_mode", "mp_divergence",
    "mp_kernel_timing", "mp_kernel_timing_get", "mp_guess_param_count", "mp_guess_train",
    "mp_guess_params", "mp_guess_last_loss", "mp_guess_eval", "mp_precond_info",
    "mp_graph_mode", "mp_batch_upload", "mp_batch_params", "mp_batch_stats", "mp_autotune",
    "mp_operator_eigenvalue", "mp_knn_stencils", "mp_rcm_order", "mp_fs_body_force",
    "mp_divergence_noalloc", "mp_precond_apply", "mp_guess_order", "mp_guess_window",
    "mp_guess_base",
)


class MemphysError(RuntimeError):
    pass


class FsParams(C.Structure):
    _fields_ = [
        ("rho", C.c_double), ("mu", C.c_double), ("dt", C.c_double),
        ("temporal_order", C.c_int), ("solver_tolerance", C.c_double),
        ("n_iter", C.c_int), ("precision", C.c_int), ("inner_tolerance", C.c_double),
        ("check_every", C.c_int), ("preconditioner", C.c_int),
        ("krylov", C.c_int), ("restart", C.c_int),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("outer_cycles", C.c_int), ("restarts", C.c_int),
        ("converged", C.c_int), ("rel_res", C.c_double), ("abs_res", C.c_double),
        ("alpha", C.c_double), ("guess_rel_res", C.c_double),
        ("total_iterations", C.c_long), ("total_solves", C.c_long),
        ("gmres_switches", C.c_int), ("guess_rejected", C.c_int),
        ("guess_skipped", C.c_int),
    ]


_hip = None
_host = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def hip_lib() -> C.CDLL:
    """Loads libmemphys_hip.so (raises if it has not been built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise MemphysError(f"{HIP_LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(HIP_LIB_PATH)
        vp = C.c_void_p
        lib.mp_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        lib.mp_ctx_destroy.argtypes = [vp]
        lib.mp_last_error.argtypes = [vp]
        lib.mp_last_error.restype = C.c_char_p
        lib.mp_sync.argtypes = [vp]
        lib.mp_cloud_upload.argtypes = [vp, C.c_int, C.c_int, _dp, _u8, _dp, _ip, _ip]
        lib.mp_weights_assemble.argtypes = [vp, C.c_int, C.c_int, _ip, C.c_int, vp, vp, vp, vp, vp]
        lib.mp_weights_set.argtypes = [vp, _dp, _dp, vp, _dp]
        lib.mp_operator_apply.argtypes = [vp, C.c_int, _dp, _dp]
        lib.mp_solver_init.argtypes = [vp, _u8, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_int, C.c_int]
        lib.mp_solver_system_size.argtypes = [vp, C.POINTER(C.c_int)]
        lib.mp_solver_solve.argtypes = [vp, _dp, _dp, _dp]
        lib.mp_solver_krylov.argtypes = [vp, C.c_int, C.c_int]
        lib.mp_fs_create.argtypes = [vp, C.POINTER(FsParams), _u8, _u8, vp, _u8]
        lib.mp_fields_set.argtypes = [vp, C.c_int, _dp]
        lib.mp_fields_get.argtypes = [vp, C.c_int, _dp]
        lib.mp_fs_body_force.argtypes = [vp, vp, vp, vp]
        lib.mp_fs_step.argtypes = [vp, C.c_int, C.c_int]
        lib.mp_stats_get.argtypes = [vp, C.POINTER(StepStats)]
        lib.mp_guess_mode.argtypes = [vp, C.c_int]
        lib.mp_divergence.argtypes = [vp, _dp]
        lib.mp_divergence_noalloc.argtypes = [vp, vp]
        lib.mp_kernel_timing.argtypes = [vp, C.c_int]
        lib.mp_kernel_timing_get.argtypes = [vp, C.c_int, C.POINTER(C.c_long),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.POINTER(C.c_long)]
        lib.mp_guess_param_count.argtypes = []
        lib.mp_guess_train.argtypes = [vp, C.c_int, C.c_double, C.c_ulonglong]
        lib.mp_guess_params.argtypes = [vp, C.c_int, np.ctypeslib.ndpointer(dtype=np.float32,
                                                                            flags="C_CONTIGUOUS")]
        lib.mp_guess_last_loss.argtypes = [vp, C.POINTER(C.c_double)]
        lib.mp_guess_eval.argtypes = [vp, _dp, _dp, _dp, vp, _dp, vp, C.POINTER(C.c_double)]
        lib.mp_graph_mode.argtypes = [vp, C.c_int]
        lib.mp_batch_upload.argtypes = [vp, C.c_int, _ip, C.c_int, _dp, _u8, _dp, _ip, _ip]
        lib.mp_batch_params.argtypes = [vp, _dp, _dp, _dp]
        lib.mp_batch_stats.argtypes = [vp, C.c_int, C.POINTER(StepStats)]
        lib.mp_autotune.argtypes = [vp, C.POINTER(C.c_int), _dp]
        lib.mp_operator_eigenvalue.argtypes = [vp, C.c_int, C.c_double, C.c_int,
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.mp_precond_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int), vp, vp,
                                        C.POINTER(C.c_double)]
        lib.mp_precond_apply.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        lib.mp_guess_order.argtypes = [vp, C.c_int]
        lib.mp_guess_window.argtypes = [vp, C.c_int]
        lib.mp_guess_base.argtypes = [vp, C.c_int]
        lib.mp_knn_stencils.argtypes = [vp, C.c_int, C.c_int, _dp, _u8, _ip, _ip, _ip]
        lib.mp_rcm_order.argtypes = [vp, C.c_int, _ip, _ip, _ip]
        for name in HIP_SYMBOLS:
            getattr(lib, name).restype = C.c_char_p if name == "mp_last_error" else C.c_int
        _hip = lib
    return _hip


def host_lib() -> C.CDLL:
    """Loads libmemphys_host.so (cloud generator, kNN, RCM)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise MemphysError(f"{HOST_LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(HOST_LIB_PATH)
        lib.mph_box_cloud_count.argtypes = [C.c_int, C.c_int]
        lib.mph_box_cloud_fill.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, _ip,
                                           _dp, _u8, _dp, _ip, _ip]
        lib.mph_rng_uniform.argtypes = [C.c_uint64, C.c_int, C.c_void_p, C.c_void_p]
        lib.mph_knn.argtypes = [C.c_int, C.c_int, _dp, _u8, _ip, _ip, _ip, C.c_int]
        lib.mph_rcm.argtypes = [C.c_int, _ip, _ip, _ip]
        lib.mph_apply_order.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, _u8, C.c_void_p, _ip, _ip]
        lib.mph_poly_exponents.argtypes = [C.c_int, C.c_int, _ip]
        lib.mph_amg_first_level.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, C.c_double,
                                            C.c_double, C.c_double, _ip,
                                            C.POINTER(C.c_int), _ip, _ip, _dp, C.c_long, _ip,
                                            _ip, _dp, C.c_long]
        lib.mph_amg_bordered_inverse.argtypes = [C.c_int, _ip, _ip, _dp, _dp]
        lib.mph_amg_dense_inverse.argtypes = [C.c_int, _ip, _ip, _dp, _dp]
        lib.mph_hessenberg_eig.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp]
        lib.mph_gmsh_read.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _host = lib
    return _host


# ----------------------------------------------------------------------------- host set-up
@dataclass
class Cloud:
    """A point cloud with its stencils (CSR, entry 0 of each row is the node)."""
    dim: int
    xyz: np.ndarray       # (nv, dim)
    boundary: np.ndarray  # (nv,) uint8
    normal: np.ndarray    # (nv, dim)
    row_ptr: np.ndarray   # (nv+1,) int32
    col: np.ndarray       # (nnz,) int32
    ksize: np.ndarray     # (nv,) int32
    layer: np.ndarray     # (nv,) int32

    @property
    def nv(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])


def box_cloud(dim: int, n: int, jitter: float, seed: int, k_interior: int,
              k_layers=(None, None, None), knn_threads: int = 0, rcm: bool = True) -> Cloud:
    """Seeded jittered box cloud, kNN stencils and RCM order (CLOUD's set-up)."""
    lib = host_lib()
    nv = lib.mph_box_cloud_count(dim, n)
    if nv <= 0:
        raise MemphysError("box_cloud: bad dim/n")
    kl = np.array([k if k is not None else k_interior for k in k_layers], dtype=np.int32)
    xyz = np.zeros(nv * dim)
    nrm = np.zeros(nv * dim)
    bnd = np.zeros(nv, dtype=np.uint8)
    ks = np.zeros(nv, dtype=np.int32)
    layer = np.zeros(nv, dtype=np.int32)
    got = lib.mph_box_cloud_fill(dim, n, float(jitter), int(seed), int(k_interior), kl, xyz, bnd,
                                 nrm, ks, layer)
    if got != nv:
        raise MemphysError("box_cloud: generator failed")
    rp = np.zeros(nv + 1, dtype=np.int32)
    col = np.zeros(int(ks.sum()), dtype=np.int32)
    if lib.mph_knn(nv, dim, xyz, bnd, ks, rp, col, int(knn_threads)) != 0:
        raise MemphysError("box_cloud: kNN failed")
    if rcm:
        order = np.zeros(nv, dtype=np.int32)
        if lib.mph_rcm(nv, rp, col, order) != 0:
            raise MemphysError("box_cloud: RCM failed")
        extra = layer.copy()
        if lib.mph_apply_order(nv, dim, order, xyz, nrm, bnd, extra.ctypes.data_as(C.c_void_p),
                               rp, col) != 0:
            raise MemphysError("box_cloud: reorder failed")
        layer = extra
        ks = np.diff(rp).astype(np.int32)
    return Cloud(dim, xyz.reshape(nv, dim), bnd, nrm.reshape(nv, dim), rp, col, ks, layer)


def gmsh_cloud(path: str, k: int, knn_threads: int = 0, rcm: bool = True) -> Cloud:
    """A Gmsh 2.2 mesh read as POINTS(PARAMETERS&) reads it (boundary flags,
    averaged outward normals, corner/edge and hanging nodes deleted), with k-node
    stencils (CLOUD::calc_cloud_points_fast) and, optionally, RCM order."""
    lib = host_lib()
    dim, nv = C.c_int(), C.c_int()
    if lib.mph_gmsh_read(path.encode(), C.byref(dim), C.byref(nv), None, None, None, None) != 0:
        raise MemphysError(f"gmsh_cloud: cannot read {path}")
    d, n = dim.value, nv.value
    xyz = np.zeros(n * d)
    nrm = np.zeros(n * d)
    bnd = np.zeros(n, dtype=np.uint8)
    orig = np.zeros(n, dtype=np.int32)
    lib.mph_gmsh_read(path.encode(), None, None, xyz.ctypes.data_as(C.c_void_p),
                      bnd.ctypes.data_as(C.c_void_p), nrm.ctypes.data_as(C.c_void_p),
                      orig.ctypes.data_as(C.c_void_p))
    ks = np.full(n, k, dtype=np.int32)
    rp = np.zeros(n + 1, dtype=np.int32)
    col = np.zeros(n * k, dtype=np.int32)
    if lib.mph_knn(n, d, xyz, bnd, ks, rp, col, int(knn_threads)) != 0:
        raise MemphysError("gmsh_cloud: kNN failed")
    extra = orig.copy()
    if rcm:
        order = np.zeros(n, dtype=np.int32)
        if lib.mph_rcm(n, rp, col, order) != 0:
            raise MemphysError("gmsh_cloud: RCM failed")
        if lib.mph_apply_order(n, d, order, xyz, nrm, bnd, extra.ctypes.data_as(C.c_void_p),
                               rp, col) != 0:
            raise MemphysError("gmsh_cloud: reorder failed")
    return Cloud(d, xyz.reshape(n, d), bnd, nrm.reshape(n, d), rp, col, ks, extra)


def poly_exponents(dim: int, poly_deg: int) -> np.ndarray:
    m = (poly_deg + 1) * (poly_deg + 2) // 2 if dim == 2 else \
        (poly_deg + 1) * (poly_deg + 2) * (poly_deg + 3) // 6
    out = np.zeros(m * dim, dtype=np.int32)
    if host_lib().mph_poly_exponents(dim, poly_deg, out) != m:
        raise MemphysError("poly_exponents failed")
    return out.reshape(m, dim)


# ----------------------------------------------------------------------------- GPU context
class Context:
    """One HIP context of the C ABI (one per host thread)."""

    # Default of new contexts: replay Krylov chunks from hipGraphs (False
    # launches every kernel; results are identical).
    default_graphs = True

    def __init__(self, device: int = 0):
        self.lib = hip_lib()
        self.h = C.c_void_p()
        rc = self.lib.mp_ctx_create(device, C.byref(self.h))
        if rc != 0:
            raise MemphysError(f"mp_ctx_create failed with status {rc}")
        self.cloud = None
        self.n_sys = None
        if not Context.default_graphs:
            self.graph_mode(False)

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.mp_last_error(self.h)
            raise MemphysError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.mp_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, cloud: Cloud):
        self.cloud = cloud
        self._check(self.lib.mp_cloud_upload(
            self.h, cloud.nv, cloud.dim, np.ascontiguousarray(cloud.xyz.ravel()),
            np.ascontiguousarray(cloud.boundary), np.ascontiguousarray(cloud.normal.ravel()),
            np.ascontiguousarray(cloud.row_ptr), np.ascontiguousarray(cloud.col)),
            "mp_cloud_upload")

    def upload_batch(self, clouds):
        """Uploads independent clouds as one batch (concatenated in order)."""
        cat = concat_clouds(clouds)
        sizes = np.ascontiguousarray([c.nv for c in clouds], dtype=np.int32)
        self.cloud = cat
        self.batch_sizes = sizes
        self._check(self.lib.mp_batch_upload(
            self.h, len(clouds), sizes, cat.dim, np.ascontiguousarray(cat.xyz.ravel()),
            np.ascontiguousarray(cat.boundary), np.ascontiguousarray(cat.normal.ravel()),
            np.ascontiguousarray(cat.row_ptr), np.ascontiguousarray(cat.col)), "mp_batch_upload")
        return cat

    def batch_params(self, rho, mu, dt):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self._check(self.lib.mp_batch_params(self.h, f(rho), f(mu), f(dt)), "mp_batch_params")

    def eigenvalue(self, op: int, tol: float = 1e-12, max_restarts: int = 200) -> complex:
        """Largest-magnitude eigenvalue of a stencil operator (imaginary part >= 0)."""
        re, im = C.c_double(), C.c_double()
        self._check(self.lib.mp_operator_eigenvalue(self.h, op, tol, max_restarts, C.byref(re),
                                                    C.byref(im)), "mp_operator_eigenvalue")
        return complex(re.value, im.value)

    def knn_stencils(self, xyz: np.ndarray, boundary: np.ndarray, k_per_node: np.ndarray):
        """Stencils on the GPU (CLOUD::calc_cloud_points_fast): returns (row_ptr, col)."""
        nv, dim = xyz.shape
        ks = np.ascontiguousarray(k_per_node, dtype=np.int32)
        rp = np.zeros(nv + 1, dtype=np.int32)
        col = np.zeros(int(ks.sum()), dtype=np.int32)
        self._check(self.lib.mp_knn_stencils(self.h, nv, dim,
                                             np.ascontiguousarray(xyz, dtype=np.float64).ravel(),
                                             np.ascontiguousarray(boundary, dtype=np.uint8), ks,
                                             rp, col), "mp_knn_stencils")
        return rp, col

    def rcm_order(self, row_ptr: np.ndarray, col: np.ndarray) -> np.ndarray:
        """Reverse Cuthill-McKee order on the GPU: order[old] = new."""
        nv = len(row_ptr) - 1
        order = np.zeros(nv, dtype=np.int32)
        self._check(self.lib.mp_rcm_order(self.h, nv,
                                          np.ascontiguousarray(row_ptr, dtype=np.int32),
                                          np.ascontiguousarray(col, dtype=np.int32), order),
                    "mp_rcm_order")
        return order

    def autotune(self):
        """Picks the fastest convergence-poll interval on the last step's solve;
        returns (interval, {interval: ms})."""
        chosen = C.c_int()
        ms = np.zeros(5)
        self._check(self.lib.mp_autotune(self.h, C.byref(chosen), ms), "mp_autotune")
        return chosen.value, dict(zip((4, 8, 16, 32, 64), ms.tolist()))

    def batch_stats(self, problem: int) -> StepStats:
        st = StepStats()
        self._check(self.lib.mp_batch_stats(self.h, problem, C.byref(st)), "mp_batch_stats")
        return st

    def weights(self, phs_deg: int, poly_deg: int, precision: int = MP_PREC_FP64,
                download: bool = True, cond: bool = False):
        c = self.cloud
        ex = np.ascontiguousarray(poly_exponents(c.dim, poly_deg).ravel())
        outs = [np.zeros(c.nnz) if download else None for _ in range(4)]
        if c.dim == 2 and download:
            outs[2] = None
        cnd = np.zeros(c.nv) if cond else None
        ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        self._check(self.lib.mp_weights_assemble(self.h, phs_deg, poly_deg, ex, precision,
                                                 ptr(outs[0]), ptr(outs[1]), ptr(outs[2]),
                                                 ptr(outs[3]), ptr(cnd)), "mp_weights_assemble")
        return outs, cnd

    def set_weights(self, gx, gy, gz, lap):
        gzp = np.ascontiguousarray(gz).ctypes.data_as(C.c_void_p) if gz is not None else None
        self._check(self.lib.mp_weights_set(self.h, np.ascontiguousarray(gx),
                                            np.ascontiguousarray(gy), gzp,
                                            np.ascontiguousarray(lap)), "mp_weights_set")

    def apply(self, op: int, x: np.ndarray) -> np.ndarray:
        y = np.zeros(self.cloud.nv)
        self._check(self.lib.mp_operator_apply(self.h, op, np.ascontiguousarray(x, dtype=np.float64),
                                               y), "mp_operator_apply")
        return y

    def solver_init(self, dirichlet, unsteady=0.0, conv=0.0, diff=1.0, tol=1e-10, n_iter=1000,
                    precision=MP_PREC_FP64):
        self._check(self.lib.mp_solver_init(self.h, np.ascontiguousarray(dirichlet, dtype=np.uint8),
                                            unsteady, conv, diff, tol, n_iter, precision),
                    "mp_solver_init")
        n = C.c_int()
        self._check(self.lib.mp_solver_system_size(self.h, C.byref(n)), "mp_solver_system_size")
        self.n_sys = n.value
        return n.value

    def solver_krylov(self, krylov, restart=0):
        """Krylov method of the solver made by solver_init (MP_KRYLOV_*)."""
        self._check(self.lib.mp_solver_krylov(self.h, krylov, restart), "mp_solver_krylov")

    def solve(self, rhs, guess):
        xn = np.zeros(self.n_sys)
        xo = np.ascontiguousarray(guess, dtype=np.float64).copy()
        self._check(self.lib.mp_solver_solve(self.h, np.ascontiguousarray(rhs, dtype=np.float64),
                                             xn, xo), "mp_solver_solve")
        return xn, xo

    def fs_create(self, rho, mu, dt, order, tol, n_iter, u_dir, v_dir, w_dir, p_dir,
                  precision=MP_PREC_FP64, inner_tol=1e-3, check_every=8,
                  precond=MP_PRECOND_JACOBI, krylov=MP_KRYLOV_BICGSTAB, restart=0):
        p = FsParams(rho, mu, dt, order, tol, n_iter, precision, inner_tol, check_every, precond,
                     krylov, restart)
        wd = np.ascontiguousarray(w_dir, dtype=np.uint8) if w_dir is not None else None
        self._wd = wd
        self._check(self.lib.mp_fs_create(
            self.h, C.byref(p), np.ascontiguousarray(u_dir, dtype=np.uint8),
            np.ascontiguousarray(v_dir, dtype=np.uint8),
            wd.ctypes.data_as(C.c_void_p) if wd is not None else None,
            np.ascontiguousarray(p_dir, dtype=np.uint8)), "mp_fs_create")
        n = C.c_int()
        self._check(self.lib.mp_solver_system_size(self.h, C.byref(n)), "mp_solver_system_size")
        self.n_sys = n.value

    def set_field(self, which: int, values):
        self._check(self.lib.mp_fields_set(self.h, which, np.ascontiguousarray(values,
                                                                               dtype=np.float64)),
                    "mp_fields_set")

    def get_field(self, which: int) -> np.ndarray:
        n = self.n_sys if which == MP_P else self.cloud.nv
        out = np.zeros(n)
        self._check(self.lib.mp_fields_get(self.h, which, out), "mp_fields_get")
        return out

    def body_force(self, fx=None, fy=None, fz=None):
        """Body force of the following steps (nv entries each); no arguments
        removes it."""
        self._force = [np.ascontiguousarray(f, dtype=np.float64) if f is not None else None
                       for f in (fx, fy, fz)]
        ptrs = [f.ctypes.data_as(C.c_void_p) if f is not None else None for f in self._force]
        self._check(self.lib.mp_fs_body_force(self.h, *ptrs), "mp_fs_body_force")

    def step(self, n_steps: int, it0: int):
        self._check(self.lib.mp_fs_step(self.h, n_steps, it0), "mp_fs_step")

    def stats(self) -> StepStats:
        s = StepStats()
        self._check(self.lib.mp_stats_get(self.h, C.byref(s)), "mp_stats_get")
        return s

    def divergence_noalloc(self, out_dev: int):
        """mp_divergence_noalloc: the 6 statistics into device memory at out_dev
        (an integer device address), stream-ordered, without allocating."""
        self._check(self.lib.mp_divergence_noalloc(self.h, C.c_void_p(out_dev)),
                    "mp_divergence_noalloc")

    def divergence(self) -> np.ndarray:
        out = np.zeros(6)
        self._check(self.lib.mp_divergence(self.h, out), "mp_divergence")
        return out

    def graph_mode(self, enable: bool):
        """Replay Krylov chunks from hipGraphs (default) or launch them one by one."""
        self._check(self.lib.mp_graph_mode(self.h, 1 if enable else 0), "mp_graph_mode")

    def precond_info(self) -> dict:
        """Multigrid hierarchy: rows and stored entries per level, set-up seconds."""
        lv = C.c_int()
        rows = np.zeros(16, dtype=np.int32)
        nnz = np.zeros(16, dtype=np.int64)
        secs = C.c_double()
        self._check(self.lib.mp_precond_info(self.h, 16, C.byref(lv),
                                             rows.ctypes.data_as(C.c_void_p),
                                             nnz.ctypes.data_as(C.c_void_p), C.byref(secs)),
                    "mp_precond_info")
        L = min(lv.value, 16)
        return {"levels": lv.value, "rows": rows[:L].tolist(), "nnz": nnz[:L].tolist(),
                "setup_s": secs.value}

    def precond_apply(self, x: np.ndarray, plain: bool = False) -> np.ndarray:
        """One multigrid V-cycle of the pressure solve on x (system size): the
        solver's cycle, or with plain=True the level-by-level one."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self._check(self.lib.mp_precond_apply(self.h, 1 if plain else 0, len(x), x, out),
                    "mp_precond_apply")
        return out

    def guess_mode(self, mode: int):
        self._check(self.lib.mp_guess_mode(self.h, mode), "mp_guess_mode")

    def guess_order(self, order: int):
        """Pressures (2, 3 or 4) the extrapolated guess goes through; default 4."""
        self._check(self.lib.mp_guess_order(self.h, order), "mp_guess_order")

    def guess_window(self, k: int):
        """Solutions (1..24) the projected guess combines; default 20."""
        self._check(self.lib.mp_guess_window(self.h, k), "mp_guess_window")

    def guess_base(self, base: int):
        """The classical guess the learned network corrects (and falls back to):
        MP_GUESS_PROJECT (default) or MP_GUESS_EXTRAPOLATE."""
        self._check(self.lib.mp_guess_base(self.h, base), "mp_guess_base")

    def guess_train(self, enable: bool, lr: float = 1e-3, seed: int = 0):
        self._check(self.lib.mp_guess_train(self.h, 1 if enable else 0, lr, seed), "mp_guess_train")

    def guess_params(self, values=None) -> np.ndarray:
        n = self.lib.mp_guess_param_count()
        if values is not None:
            v = np.ascontiguousarray(values, dtype=np.float32)
            assert v.size == n
            self._check(self.lib.mp_guess_params(self.h, 1, v), "mp_guess_params")
            return v
        out = np.zeros(n, dtype=np.float32)
        self._check(self.lib.mp_guess_params(self.h, 0, out), "mp_guess_params")
        return out

    def guess_last_loss(self) -> float:
        v = C.c_double()
        self._check(self.lib.mp_guess_last_loss(self.h, C.byref(v)), "mp_guess_last_loss")
        return v.value

    def guess_eval(self, p1, p2, b, p_true=None):
        nv = self.cloud.nv
        g = np.zeros(nv)
        grad = np.zeros(self.lib.mp_guess_param_count(), dtype=np.float32) if p_true is not None \
            else None
        loss = C.c_double()
        pt = np.ascontiguousarray(p_true, dtype=np.float64) if p_true is not None else None
        self._check(self.lib.mp_guess_eval(
            self.h, np.ascontiguousarray(p1, dtype=np.float64),
            np.ascontiguousarray(p2, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64),
            pt.ctypes.data_as(C.c_void_p) if pt is not None else None, g,
            grad.ctypes.data_as(C.c_void_p) if grad is not None else None, C.byref(loss)),
            "mp_guess_eval")
        return g, grad, loss.value

    def sync(self):
        self._check(self.lib.mp_sync(self.h), "mp_sync")

    def kernel_timing(self, enable: bool):
        self._check(self.lib.mp_kernel_timing(self.h, 1 if enable else 0), "mp_kernel_timing")

    def kernel_timing_get(self, which: int):
        """(working launches, their mean ms, algorithmic bytes per launch, all launches):
        launches of a converged system, which return at once, are not working ones."""
        n = C.c_long()
        ms = C.c_double()
        b = C.c_double()
        a = C.c_long()
        self._check(self.lib.mp_kernel_timing_get(self.h, which, C.byref(n), C.byref(ms),
                                                  C.byref(b), C.byref(a)),
                    "mp_kernel_timing_get")
        return n.value, ms.value, b.value, a.value


# ----------------------------------------------------------------------------- learned guess
def lid_conditions(cloud: Cloud, lid_speed: float = 1.0):
    """Cavity boundary conditions: velocity Dirichlet on every wall, the lid
    (last coordinate = 1) moving along x at lid_speed; pressure Neumann."""
    nv = cloud.nv
    u_dir = np.ones(nv, dtype=np.uint8)
    p_dir = np.zeros(nv, dtype=np.uint8)
    lid = (cloud.boundary == 1) & (np.abs(cloud.xyz[:, cloud.dim - 1] - 1.0) < 1e-12)
    u0 = np.where(lid, lid_speed, 0.0)
    return u_dir, p_dir, u0


def train_guess(cloud: Cloud, rho: float, mu: float, dt: float, tol: float, episodes: int,
                steps_per_episode: int, lid_range=(0.5, 1.5), seed: int = 1, lr: float = 1e-3,
                device: int = 0, log=None, **solver) -> np.ndarray:
    """Trains the learned pressure guess on solver-generated pairs: `episodes`
    cavity runs from rest with seeded random lid speeds in lid_range, one Adam
    step per pressure solve (forward/backward/Adam run in the HIP library).
    `solver`: fs_create's solver options (precision, precond, krylov,
    inner_tol); the targets, converged pressures, do not depend on them.
    Returns the packed FP32 parameters."""
    rng = np.random.default_rng(seed)
    with Context(device) as ctx:
        ctx.upload(cloud)
        ctx.weights(3, 3, download=False)
        u_dir, p_dir, u0 = lid_conditions(cloud)
        w_dir = u_dir if cloud.dim == 3 else None
        ctx.fs_create(rho, mu, dt, 2, tol, 20000, u_dir, u_dir, w_dir, p_dir, **solver)
        ctx.guess_mode(MP_GUESS_LEARNED)
        ctx.guess_train(True, lr, seed)
        nv = cloud.nv
        for ep in range(episodes):
            speed = float(rng.uniform(*lid_range))
            ctx.set_field(MP_U, u0 * speed)
            ctx.set_field(MP_V, np.zeros(nv))
            if cloud.dim == 3:
                ctx.set_field(MP_W, np.zeros(nv))
            ctx.set_field(MP_P, np.zeros(ctx.n_sys))
            iters = 0
            for it in range(steps_per_episode):
                ctx.step(1, it)
                iters += ctx.stats().iterations
            if log:
                log(f"guess training episode {ep}: lid {speed:.3f}, mean iterations "
                    f"{iters / steps_per_episode:.1f}, last loss {ctx.guess_last_loss():.3e}")
        return ctx.guess_params()


def rng_uniform(seed: int, count: int) -> np.ndarray:
    """First `count` values of the library's seeded generator (std::mt19937_64,
    53-bit mapping (x >> 11) * 2^-53)."""
    out = np.zeros(count)
    host_lib().mph_rng_uniform(seed, count, out.ctypes.data_as(C.c_void_p), None)
    return out


def cavity_batch(nprob: int = 64, seed: int = 5, n_min: int = 2000, n_max: int = 40000,
                 re_range=(50.0, 400.0), k: int = 30, jitter: float = 0.25):
    """The batched-cavities workload: nprob 2D lid-driven cavities with node
    counts log-uniform on [n_min, n_max] (lattice n x n minus the 4 corners),
    Reynolds numbers uniform on re_range (rho = 1, mu = 1/Re, lid speed 1),
    interior jitter +-jitter*h, PHS r^3 + degree-3 polynomial, k-node stencils.
    The time step of each problem is half its explicit AB2 limit,
    dt = min(0.5 h^2 Re / 33.7, 0.2 h) (33.7 h^-2 bounds |lambda(L)| on these
    clouds).  Returns [(cloud, Re, dt)]."""
    u = rng_uniform(seed, 2 * nprob)
    out = []
    for q in range(nprob):
        nodes = np.exp(np.log(n_min) + u[2 * q] * (np.log(n_max) - np.log(n_min)))
        n = max(8, int(round(np.sqrt(nodes + 4.0))))
        re = re_range[0] + u[2 * q + 1] * (re_range[1] - re_range[0])
        h = 1.0 / (n - 1)
        dt = min(0.5 * h * h * re / 33.7, 0.2 * h)
        cloud = box_cloud(dim=2, n=n, jitter=jitter, seed=seed * 1000 + q, k_interior=k)
        out.append((cloud, re, dt))
    return out


def concat_clouds(clouds) -> "Cloud":
    """Disjoint union of clouds (node indices of each shifted past the earlier ones)."""
    dim = clouds[0].dim
    if any(c.dim != dim for c in clouds):
        raise MemphysError("concat_clouds: mixed dimensions")
    offs = np.cumsum([0] + [c.nv for c in clouds])
    nnz_offs = np.cumsum([0] + [c.nnz for c in clouds])
    row_ptr = np.concatenate([c.row_ptr[:-1] + nnz_offs[k] for k, c in enumerate(clouds)] +
                             [np.array([nnz_offs[-1]])]).astype(np.int32)
    col = np.concatenate([c.col + offs[k] for k, c in enumerate(clouds)]).astype(np.int32)
    return Cloud(dim=dim, xyz=np.concatenate([c.xyz for c in clouds]),
                 boundary=np.concatenate([c.boundary for c in clouds]),
                 normal=np.concatenate([c.normal for c in clouds]), row_ptr=row_ptr, col=col,
                 ksize=np.concatenate([c.ksize for c in clouds]),
                 layer=np.concatenate([c.layer for c in clouds]))


def amg_first_level(A, k: int = 6, omega_p: float = 0.66, trunc: float = 0.0,
                    drop: float = 0.0):
    """First coarsening step of the library's multigrid set-up on a scipy CSR
    matrix: (aggregate per row, P, R A P) as computed by the C++ builder, with
    P truncated at `trunc` and R A P sparsified at `drop` (the library's
    hierarchy uses 0.15 and 0.01; 0 turns either off)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    col = np.ascontiguousarray(A.indices, dtype=np.int32)
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    agg = np.zeros(n, dtype=np.int32)
    na = C.c_int()
    pcap = int(A.nnz + n)
    p_ptr = np.zeros(n + 1, dtype=np.int32)
    p_col = np.zeros(pcap, dtype=np.int32)
    p_val = np.zeros(pcap)
    ac_cap = int(8 * A.nnz)
    ac_ptr = np.zeros(n + 1, dtype=np.int32)
    ac_col = np.zeros(ac_cap, dtype=np.int32)
    ac_val = np.zeros(ac_cap)
    rc = host_lib().mph_amg_first_level(n, ptr, col, val, k, omega_p, trunc, drop, agg,
                                        C.byref(na), p_ptr, p_col, p_val, pcap, ac_ptr, ac_col,
                                        ac_val, ac_cap)
    if rc != 0:
        raise MemphysError(f"mph_amg_first_level failed ({rc})")
    m = na.value
    P = sp.csr_matrix((p_val[:p_ptr[n]], p_col[:p_ptr[n]], p_ptr), shape=(n, m))
    Ac = sp.csr_matrix((ac_val[:ac_ptr[m]], ac_col[:ac_ptr[m]], ac_ptr[:m + 1]), shape=(m, m))
    return agg, P, Ac


def amg_bordered_inverse(A) -> np.ndarray:
    """The library's coarsest-level solve: top-left block of inv([[A, 1], [1^T, 0]])."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    G = np.zeros(n * n)
    rc = host_lib().mph_amg_bordered_inverse(n, np.ascontiguousarray(A.indptr, dtype=np.int32),
                                             np.ascontiguousarray(A.indices, dtype=np.int32),
                                             np.ascontiguousarray(A.data, dtype=np.float64), G)
    if rc != 0:
        raise MemphysError("mph_amg_bordered_inverse failed")
    return G.reshape(n, n)


def amg_dense_inverse(A) -> np.ndarray:
    """The library's coarsest-level solve of a system with Dirichlet rows: inv(A)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    G = np.zeros(n * n)
    rc = host_lib().mph_amg_dense_inverse(n, np.ascontiguousarray(A.indptr, dtype=np.int32),
                                          np.ascontiguousarray(A.indices, dtype=np.int32),
                                          np.ascontiguousarray(A.data, dtype=np.float64), G)
    if rc != 0:
        raise MemphysError("mph_amg_dense_inverse failed")
    return G.reshape(n, n)


def hessenberg_eig(H, which: int = 0):
    """Eigenvalues of an upper Hessenberg matrix and the unit eigenvector of
    eigenvalue number `which`, by the routines of the GPU Arnoldi iteration."""
    H = np.ascontiguousarray(H, dtype=np.float64)
    m = H.shape[0]
    re, im, vr, vi = (np.zeros(m) for _ in range(4))
    if host_lib().mph_hessenberg_eig(m, H.ravel(), re, im, which, vr, vi) != 0:
        raise MemphysError("hessenberg_eig failed")
    return re + 1j * im, vr + 1j * vi
