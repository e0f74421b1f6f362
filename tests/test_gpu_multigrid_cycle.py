"""The multigrid V(1,1)-cycle that preconditions the pressure solve, applied on
its own through mp_precond_apply: with three or more levels the solver's cycle
runs the levels below level 1 from composed matrices (Q_l = R_{l-1} - A_l D_l
R_{l-1}, T_l = P_l - D_l A_l P_l, and the coarsest solve times the restriction
to it), which is the same linear map as the level-by-level cycle; only the
BF16 rounding of the stored values differs (one rounding of each composed
entry instead of one per factor).  A wrong sign, a dropped term or a wrong
level offset in the composed steps moves the result by O(1); the rounding
moves it by about 1e-3."""
import numpy as np
import pytest

from conftest import lid_bc, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import memphys_amd as M
    c = M.Context(0)
    yield c
    c.close()


def _print(*a):
    print(*a, flush=True)


def _check_cycles(ctx, nsys, nv, bordered, seed):
    rng = np.random.default_rng(seed)
    for trial in range(2):
        x = rng.standard_normal(nsys)
        yc = ctx.precond_apply(x)
        yp = ctx.precond_apply(x, plain=True)
        e = rel_err(yc, yp)
        _print(f"  vector {trial}: composed vs level-by-level {e:.2e}")
        assert np.all(np.isfinite(yc))
        assert e <= 1e-2
        assert not np.array_equal(yc, yp)  # the composed steps ran
        if bordered:
            assert np.array_equal(yc[nv:], x[nv:])  # border unknowns pass through
        # the same map: linear in its input, to FP32 rounding
        y2 = ctx.precond_apply(2.0 * x - 0.5 * yc)
        lin = 2.0 * yc - 0.5 * ctx.precond_apply(yc)
        assert rel_err(y2, lin) <= 1e-4


def _fs(ctx, cloud, p_dir):
    import memphys_amd as M
    ctx.upload(cloud)
    ctx.weights(3, 3, download=False)
    u_dir, _, _ = lid_bc(cloud, cloud.dim - 1)
    ctx.fs_create(1.0, 0.01, 1e-4, 2, 1e-10, 2000, u_dir, u_dir,
                  u_dir if cloud.dim == 3 else None, p_dir, precision=M.MP_PREC_MIXED,
                  precond=M.MP_PRECOND_AMG, krylov=M.MP_KRYLOV_GMRES)
    return ctx.precond_info()


def test_composed_cycle_2d_three_levels(ctx, cavity2d):
    """10k-node 2D cavity, full-Neumann pressure (bordered system)."""
    info = _fs(ctx, cavity2d, np.zeros(cavity2d.nv, dtype=np.uint8))
    _print("2D cavity hierarchy", info)
    assert info["levels"] >= 3
    _check_cycles(ctx, ctx.n_sys, cavity2d.nv, True, 11)


def test_composed_cycle_3d_four_levels(ctx):
    """175k-node 3D box with the 1M-node case's stencil rule and a Dirichlet
    pressure face (plain coarsest inverse, no border unknown): four levels, so
    every composed step (Q_2, the coarsest solve times R_2, T_2, T_1) runs."""
    import memphys_amd as M
    cloud = M.box_cloud(dim=3, n=56, jitter=0.3, seed=7, k_interior=40, k_layers=(60, 55, 50))
    p_dir = ((cloud.boundary == 1) & (np.abs(cloud.xyz[:, 0]) < 1e-12)).astype(np.uint8)
    assert p_dir.any()
    info = _fs(ctx, cloud, p_dir)
    _print("3D box hierarchy", info)
    assert info["levels"] >= 4
    assert ctx.n_sys == cloud.nv
    _check_cycles(ctx, ctx.n_sys, cloud.nv, False, 12)


def test_composed_cycle_batch(ctx):
    """Two 2D cavities in one batch: one hierarchy per cavity stored
    block-diagonally, so the composed matrices and the coarsest rows are
    per-cavity blocks at the cavities' level offsets."""
    import memphys_amd as M
    clouds = [M.box_cloud(dim=2, n=n, jitter=0.25, seed=s, k_interior=30)
              for n, s in ((96, 3), (110, 4))]
    cat = ctx.upload_batch(clouds)
    ctx.weights(3, 3, download=False)
    u_dir, p_dir, _ = M.lid_conditions(cat)
    ctx.fs_create(1.0, 0.01, 1e-4, 2, 1e-10, 2000, u_dir, u_dir, None, p_dir,
                  precision=M.MP_PREC_MIXED, precond=M.MP_PRECOND_AMG,
                  krylov=M.MP_KRYLOV_GMRES)
    info = ctx.precond_info()
    _print("batch hierarchy", info)
    assert info["levels"] >= 3
    _check_cycles(ctx, ctx.n_sys, cat.nv, True, 13)


def test_autotune_leaves_fields_and_keeps_accuracy(ctx, cavity2d):
    """mp_autotune (poll interval by timing; multigrid layouts by timing each
    candidate after a cache-clearing read) leaves the fields and pressure
    history alone, and the steps after it match an untuned run to the
    solver's accuracy (the layouts change the V-cycle at the rounding level)."""
    import memphys_amd as M
    cloud = cavity2d
    u_dir, p_dir, u0 = lid_bc(cloud, 1)

    def run(tune):
        ctx.upload(cloud)
        ctx.weights(3, 3, download=False)
        ctx.fs_create(1.0, 0.01, 2e-4, 2, 1e-12, 2000, u_dir, u_dir, None, p_dir,
                      precision=M.MP_PREC_MIXED, precond=M.MP_PRECOND_AMG,
                      krylov=M.MP_KRYLOV_GMRES)
        ctx.set_field(M.MP_U, u0)
        ctx.set_field(M.MP_V, np.zeros(cloud.nv))
        ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
        ctx.step(3, 0)
        if tune:
            before = [ctx.get_field(f) for f in (M.MP_U, M.MP_V, M.MP_P)]
            chosen, ms = ctx.autotune()
            _print("autotune: interval", chosen, "ms", ms)
            assert chosen in (4, 8, 16, 32, 64)
            assert all(v > 0 for v in ms.values())
            after = [ctx.get_field(f) for f in (M.MP_U, M.MP_V, M.MP_P)]
            for a, b in zip(before, after):
                assert np.array_equal(a, b)
        ctx.step(3, 3)
        return [ctx.get_field(f) for f in (M.MP_U, M.MP_V, M.MP_P)]

    ref = run(False)
    tuned = run(True)
    errs = [rel_err(a, b) for a, b in zip(tuned, ref)]
    _print("tuned vs untuned after 6 steps:", errs)
    assert max(errs) <= 1e-8
