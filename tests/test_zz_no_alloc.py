"""mp_divergence_noalloc (the entry point that allocates nothing) against
mp_divergence: the same six continuity statistics, bit for bit, on two calls in
a row, the second on a smaller cloud (the shared scratch array keeps the first
call's partial sums, which the second must not read).  The output buffer is
device memory the test takes from the HIP runtime directly."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _hip():
    lib = C.CDLL("libamdhip64.so.7")  # the runtime the library is linked to
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _stepped_context(M, cloud, steps):
    ctx = M.Context(0)
    ctx.upload(cloud)
    ctx.weights(3, 3, download=False)
    u_dir, p_dir, u0 = M.lid_conditions(cloud)
    ctx.fs_create(1.0, 0.0025 if cloud.dim == 3 else 0.01, 5e-4 if cloud.dim == 3 else 2e-4, 2,
                  1e-10, 20000, u_dir, u_dir, u_dir if cloud.dim == 3 else None, p_dir)
    ctx.set_field(M.MP_U, u0)
    ctx.set_field(M.MP_V, np.zeros(cloud.nv))
    if cloud.dim == 3:
        ctx.set_field(M.MP_W, np.zeros(cloud.nv))
    ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
    ctx.step(steps, 0)
    return ctx


def test_divergence_noalloc_matches_divergence():
    import memphys_amd as M
    hip = _hip()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 6 * 8) == 0
    try:
        clouds = [M.box_cloud(dim=3, n=14, jitter=0.3, seed=7, k_interior=40,
                              k_layers=(60, 55, 50)),
                  M.box_cloud(dim=2, n=20, jitter=0.25, seed=3, k_interior=20)]
        assert clouds[1].nv < clouds[0].nv
        for cloud in clouds:
            ctx = _stepped_context(M, cloud, 3)
            ref = ctx.divergence()
            ctx.divergence_noalloc(dev.value)
            ctx.sync()
            got = np.zeros(6)
            assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), dev, 6 * 8, 2) == 0  # D2H
            _print(f"case: {cloud.dim}D nv {cloud.nv}: mp_divergence {ref.tolist()} "
                   f"noalloc {got.tolist()}")
            assert np.array_equal(got, ref)
            assert ref[4] > 0.0
            ctx.close()
    finally:
        hip.hipFree(dev)


def test_divergence_noalloc_needs_fields():
    import memphys_amd as M
    with M.Context(0) as ctx:
        with pytest.raises(M.MemphysError):
            ctx.divergence_noalloc(0)
