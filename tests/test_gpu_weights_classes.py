"""GPU stencil weights in every size class, bit for bit against the oracle.

The FP64 weights run in a register-resident kernel: one wavefront per stencil
for saddle-point systems of M = k + m <= 64 (classes M <= 32, 48, 64), two
wavefronts up to M <= 96 (classes 80, 96), and the LDS kernel beyond that and
whenever condition estimates are requested; systems too large for LDS (high
polynomial degrees) run the same kernel on a global work area.  Each class, an unjittered
lattice (exact ties in the pivot search), other PHS and polynomial degrees
and a singular stencil (two coincident nodes: a zero pivot, weights inf/NaN)
must give the oracle's weights exactly: the same roundings in the same order
(coefficient_computations.cpp:33-328, restated in oracle/oracle.cpp).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import memphys_amd as M
    c = M.Context(0)
    yield c
    c.close()


CASES = [
    # (label, box_cloud arguments, phs, poly)
    ("2D k=20, M=30", dict(dim=2, n=20, jitter=0.25, seed=3, k_interior=20), 3, 3),
    ("2D k=30, M=40", dict(dim=2, n=30, jitter=0.25, seed=4, k_interior=30), 3, 3),
    ("2D unjittered k=30", dict(dim=2, n=24, jitter=0.0, seed=4, k_interior=30), 3, 3),
    ("2D k=44 PHS5 deg4, M=59", dict(dim=2, n=22, jitter=0.25, seed=8, k_interior=44), 5, 4),
    ("3D k=40 + 60/55/50", dict(dim=3, n=10, jitter=0.3, seed=7, k_interior=40,
                               k_layers=(60, 55, 50)), 3, 3),
    ("3D unjittered k=40 + 60/55/50", dict(dim=3, n=9, jitter=0.0, seed=7, k_interior=40,
                                          k_layers=(60, 55, 50)), 3, 3),
    ("3D k=70, M=90", dict(dim=3, n=9, jitter=0.3, seed=5, k_interior=70), 3, 3),
    ("3D k=76, M=96", dict(dim=3, n=9, jitter=0.3, seed=6, k_interior=76), 3, 3),
    ("3D k=80, M=100 (LDS kernel)", dict(dim=3, n=9, jitter=0.3, seed=6, k_interior=80), 3, 3),
    ("3D k=30 PHS7 deg2, M=40", dict(dim=3, n=9, jitter=0.3, seed=2, k_interior=30), 7, 2),
    # high polynomial degrees (the reference accepts up to 15): the LDS kernel,
    # and above its LDS capacity the same kernel on a global work area
    ("3D k=70 deg4, M=105 (LDS kernel)", dict(dim=3, n=9, jitter=0.25, seed=3, k_interior=70),
     3, 4),
    ("3D k=112 deg5, M=168 (global work area)",
     dict(dim=3, n=8, jitter=0.25, seed=3, k_interior=112), 3, 5),
    ("2D k=108 deg7, M=144 (global work area)",
     dict(dim=2, n=20, jitter=0.25, seed=3, k_interior=108), 3, 7),
]


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def _check(ctx, cloud, phs, poly, label, cond=False):
    import oracle as O
    ctx.upload(cloud)
    (gx, gy, gz, lap), cnd = ctx.weights(phs, poly, cond=cond)
    ox, oy, oz, ol, oc = O.stencil_weights(cloud, phs, poly, want_cond=cond)
    got = [gx, gy, lap] + ([gz] if cloud.dim == 3 else [])
    ref = [ox, oy, ol] + ([oz] if cloud.dim == 3 else [])
    exact = [float(np.mean((g == r) | (np.isnan(g) & np.isnan(r)))) for g, r in zip(got, ref)]
    _print(f"case: {label} nv {cloud.nv} k {sorted(set(cloud.ksize.tolist()))} cond {cond}: "
           f"bit-exact fractions {exact}")
    for g, r in zip(got, ref):
        assert _same(g, r)
    if cond:
        assert _same(cnd, oc)


@pytest.mark.parametrize("label,args,phs,poly", CASES, ids=[c[0] for c in CASES])
def test_weights_bit_identical_to_oracle(ctx, label, args, phs, poly):
    import memphys_amd as M
    cloud = M.box_cloud(**args)
    _check(ctx, cloud, phs, poly, label)


def test_weights_with_condition_estimates(ctx):
    """Condition estimates requested: the LDS kernel, same weights."""
    import memphys_amd as M
    cloud = M.box_cloud(dim=3, n=8, jitter=0.3, seed=7, k_interior=40, k_layers=(60, 55, 50))
    _check(ctx, cloud, 3, 3, "3D with condition estimates", cond=True)


def test_singular_stencil_matches_oracle(ctx):
    """Two coincident nodes make every stencil holding both singular: a zero
    pivot (the oracle skips that column's elimination) and inf/NaN weights in
    those rows, identical on both sides; every other row is unaffected."""
    import memphys_amd as M
    cloud = M.box_cloud(dim=2, n=14, jitter=0.25, seed=9, k_interior=20)
    interior = np.where(cloud.boundary == 0)[0]
    i = int(interior[len(interior) // 2])
    j = int(cloud.col[cloud.row_ptr[i] + 1])  # i's nearest neighbour
    cloud.xyz[j] = cloud.xyz[i]
    _check(ctx, cloud, 3, 3, "2D with a duplicated node")
    (gx, _, _, lap), _ = ctx.weights(3, 3)
    bad = [r for r in range(cloud.nv)
           if not np.all(np.isfinite(lap[cloud.row_ptr[r]:cloud.row_ptr[r + 1]]))]
    holds_both = [r for r in range(cloud.nv)
                  if {i, j} <= set(cloud.col[cloud.row_ptr[r]:cloud.row_ptr[r + 1]].tolist())]
    _print("rows with non-finite weights", len(bad), "rows holding both nodes", len(holds_both))
    assert bad and set(bad) <= set(holds_both)


def test_mixed_weights_at_high_degree_run_in_fp64(ctx):
    """A mixed-precision request whose systems the FP32 kernel cannot take
    (degree 5 in 3D: 56 monomials, M = 168) is solved in FP64 throughout:
    bit-identical to the oracle."""
    import memphys_amd as M
    cloud = M.box_cloud(dim=3, n=8, jitter=0.25, seed=3, k_interior=112)
    ctx.upload(cloud)
    (gx, gy, gz, lap), _ = ctx.weights(3, 5, precision=M.MP_PREC_MIXED)
    import oracle as O
    ox, oy, oz, ol, _ = O.stencil_weights(cloud, 3, 5)
    for g, r in ((gx, ox), (gy, oy), (gz, oz), (lap, ol)):
        assert _same(g, r)


def test_mixed_weights_remainder_uses_same_kernel(ctx):
    """FP32 LU + FP64 refinement: the longer stencils go to the FP64 kernels and
    stay bit-identical to the oracle; the refined ones agree to 1e-12."""
    import memphys_amd as M
    import oracle as O
    cloud = M.box_cloud(dim=3, n=10, jitter=0.3, seed=7, k_interior=40, k_layers=(60, 55, 50))
    ctx.upload(cloud)
    (gx, gy, gz, lap), _ = ctx.weights(3, 3, precision=M.MP_PREC_MIXED)
    ox, oy, oz, ol, _ = O.stencil_weights(cloud, 3, 3)
    kmin = int(cloud.ksize.min())
    long_rows = np.where(cloud.ksize > kmin)[0]
    idx = np.concatenate([np.arange(cloud.row_ptr[r], cloud.row_ptr[r + 1]) for r in long_rows])
    for g, r in ((gx, ox), (gy, oy), (gz, oz), (lap, ol)):
        assert np.array_equal(g[idx], r[idx])
        assert np.linalg.norm(g - r) <= 1e-12 * np.linalg.norm(r)
