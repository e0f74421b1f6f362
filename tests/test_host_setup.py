"""Host set-up library (CPU): generator, kNN stencils, RCM, exponents, C ABI exports."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import memphys_amd as M
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mt19937_64_known_answer():
    # The C++ standard fixes the 10000th output of a default-seeded (5489)
    # std::mt19937_64: 9981545732273789042.
    raw = np.zeros(10000, dtype=np.uint64)
    M.host_lib().mph_rng_uniform(5489, 10000, None, raw.ctypes.data_as(C.c_void_p))
    assert int(raw[-1]) == 9981545732273789042
    u = np.zeros(5)
    M.host_lib().mph_rng_uniform(5489, 5, u.ctypes.data_as(C.c_void_p), None)
    assert np.all(u == (raw[:5] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)


def test_box_cloud_counts_and_geometry():
    lib = M.host_lib()
    assert lib.mph_box_cloud_count(2, 100) == 9996
    assert lib.mph_box_cloud_count(3, 100) == 998816
    c = M.box_cloud(dim=3, n=10, jitter=0.3, seed=1, k_interior=40, k_layers=(60, 55, 50))
    assert c.nv == 1000 - 12 * 8 - 8
    h = 1.0 / 9
    bnd = c.boundary == 1
    # face nodes lie on exactly one face with the outward unit normal
    on_face = (np.abs(c.xyz) < 1e-15) | (np.abs(c.xyz - 1.0) < 1e-15)
    assert np.all(on_face[bnd].sum(axis=1) == 1)
    assert np.all(on_face[~bnd].sum(axis=1) == 0)
    assert np.allclose(np.abs(c.normal[bnd]).sum(axis=1), 1.0)
    assert np.all(c.normal[~bnd] == 0.0)
    nrm_axis = np.argmax(np.abs(c.normal[bnd]), axis=1)
    side = c.xyz[bnd, :][np.arange(bnd.sum()), nrm_axis]
    assert np.all(np.sign(c.normal[bnd][np.arange(bnd.sum()), nrm_axis]) == np.where(side > 0.5, 1, -1))
    # interior jitter within +-0.3h of the lattice
    lat = np.round(c.xyz / h) * h
    assert np.abs(c.xyz - lat)[~bnd].max() <= 0.3 * h + 1e-15
    # stencil sizes by lattice layer
    ks = np.diff(c.row_ptr)
    assert set(ks[c.layer == 0]) == {60} and set(ks[c.layer == 1]) == {55}
    assert set(ks[c.layer == 2]) == {50} and set(ks[c.layer >= 3]) == {40}
    assert np.all(ks[bnd] == 60)


@pytest.mark.parametrize("dim,n,k,layers", [(2, 14, 30, (None, None, None)),
                                            (3, 8, 40, (60, 55, 50))])
def test_knn_matches_bruteforce(dim, n, k, layers):
    c = M.box_cloud(dim=dim, n=n, jitter=0.25, seed=11, k_interior=k, k_layers=layers, rcm=False)
    rp, col = O.knn_bruteforce(c.xyz, c.boundary, c.ksize)
    assert np.array_equal(rp, c.row_ptr)
    assert np.array_equal(col, c.col)
    # boundary stencils: self then interior nodes only
    for i in np.flatnonzero(c.boundary):
        row = c.col[c.row_ptr[i]:c.row_ptr[i + 1]]
        assert row[0] == i and not c.boundary[row[1:]].any()


def test_knn_rejects_oversized_stencil():
    c = M.box_cloud(dim=2, n=6, jitter=0.1, seed=1, k_interior=10, rcm=False)
    ks = np.full(c.nv, c.nv + 1, dtype=np.int32)
    rp = np.zeros(c.nv + 1, dtype=np.int32)
    col = np.zeros(int(ks.sum()), dtype=np.int32)
    xyz = np.ascontiguousarray(c.xyz.ravel())
    assert M.host_lib().mph_knn(c.nv, 2, xyz, c.boundary, ks, rp, col, 1) == -1


def test_rcm_matches_reference_restatement():
    c = M.box_cloud(dim=3, n=9, jitter=0.3, seed=5, k_interior=40, k_layers=(60, 55, 50),
                    rcm=False)
    order = np.zeros(c.nv, dtype=np.int32)
    assert M.host_lib().mph_rcm(c.nv, c.row_ptr, c.col, order) == 0
    assert np.array_equal(order, O.rcm_reference(c.row_ptr, c.col))
    assert np.array_equal(np.sort(order), np.arange(c.nv))


def test_apply_order_permutes_consistently():
    c0 = M.box_cloud(dim=2, n=12, jitter=0.25, seed=2, k_interior=30, rcm=False)
    c1 = M.box_cloud(dim=2, n=12, jitter=0.25, seed=2, k_interior=30, rcm=True)
    order = O.rcm_reference(c0.row_ptr, c0.col)
    assert np.array_equal(c1.xyz[order], c0.xyz)
    assert np.array_equal(c1.boundary[order], c0.boundary)
    for i in range(c0.nv):
        r0 = c0.col[c0.row_ptr[i]:c0.row_ptr[i + 1]]
        j = order[i]
        r1 = c1.col[c1.row_ptr[j]:c1.row_ptr[j + 1]]
        assert np.array_equal(order[r0], r1)


def test_poly_exponents_product_matches_oracle():
    for dim in (2, 3):
        for deg in range(0, 7):
            assert np.array_equal(M.poly_exponents(dim, deg), O.poly_exponents(dim, deg))


def test_hip_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "memphys_hip.h")).read()
    declared = set(re.findall(r"\b(mp_[a-z_0-9]+)\s*\(", header))
    assert declared == set(M.HIP_SYMBOLS)
    lib = M.hip_lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_host_library_exports():
    lib = M.host_lib()
    for name in ("mph_box_cloud_count", "mph_box_cloud_fill", "mph_knn", "mph_rcm",
                 "mph_apply_order", "mph_poly_exponents", "mph_rng_uniform"):
        assert hasattr(lib, name)
