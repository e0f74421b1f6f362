"""The extrapolated pressure guess of order 2, 3 and 4 (mp_guess_order): the
polynomial in time through the pressures of the last 2, 3 or 4 steps.  Every
order converges to the same fields; on a smooth run the higher orders start
closer and take fewer Krylov iterations; raising the order mid-run waits for
the history to fill (the next step still extrapolates through two pressures,
bit for bit as the order-2 run)."""
import numpy as np
import pytest

from conftest import lid_bc, rel_err

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _run(cloud, steps, order, raise_at=None, krylov=1):
    import memphys_amd as M
    u_dir, p_dir, u0 = lid_bc(cloud, cloud.dim - 1)
    out, iters = [], []
    with M.Context(0) as ctx:
        ctx.upload(cloud)
        ctx.weights(3, 3, download=False)
        ctx.fs_create(1.0, 0.0025, 5e-4, 2, 1e-10, 20000, u_dir, u_dir,
                      u_dir if cloud.dim == 3 else None, p_dir, precision=M.MP_PREC_MIXED,
                      inner_tol=1e-5, precond=M.MP_PRECOND_AMG, krylov=krylov)
        ctx.set_field(M.MP_U, u0)
        for f in (M.MP_V, M.MP_W):
            ctx.set_field(f, np.zeros(cloud.nv))
        ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
        ctx.guess_mode(M.MP_GUESS_EXTRAPOLATE)
        ctx.guess_order(order)
        with pytest.raises(M.MemphysError):
            ctx.guess_order(5)
        for it in range(steps):
            if raise_at is not None and it == raise_at:
                ctx.guess_order(4)
            ctx.step(1, it)
            st = ctx.stats()
            assert st.converged == 1 and st.rel_res <= 1e-10
            iters.append(st.iterations)
            out.append([ctx.get_field(f) for f in (M.MP_U, M.MP_V, M.MP_W, M.MP_P)])
    return out, iters


def test_orders_agree_and_higher_orders_start_closer(cube3d_small):
    steps = 40
    runs = {k: _run(cube3d_small, steps, k) for k in (2, 3, 4)}
    mean = {k: float(np.mean(runs[k][1][10:])) for k in runs}
    _print("mean GMRES iterations per step (steps 10-39) by order:", mean)
    for k in (3, 4):
        errs = [rel_err(a, b) for a, b in zip(runs[k][0][-1], runs[2][0][-1])]
        _print(f"order {k} vs 2 after {steps} steps:", errs)
        assert max(errs) <= 1e-8
    assert mean[4] < mean[2] and mean[3] < mean[2]
    # steps 0 and 1 have too short a history to extrapolate, and step 2 has
    # two pressures only: the same as order 2, bit for bit
    for it in range(3):
        for a, b in zip(runs[4][0][it], runs[2][0][it]):
            assert np.array_equal(a, b)


def test_raising_the_order_waits_for_the_history(cube3d_small):
    base, _ = _run(cube3d_small, 8, 2)
    raised, it_r = _run(cube3d_small, 8, 2, raise_at=6)
    # step 6 (the first after raising) extrapolates through two pressures still
    for it in range(7):
        for a, b in zip(raised[it], base[it]):
            assert np.array_equal(a, b)
    errs = [rel_err(a, b) for a, b in zip(raised[-1], base[-1])]
    _print("raised at step 6 vs order 2 after 8 steps:", errs, "iterations", it_r)
    assert max(errs) <= 1e-8
    assert not all(np.array_equal(a, b) for a, b in zip(raised[-1], base[-1]))


def _run_project(cloud, steps, window=16):
    import memphys_amd as M
    u_dir, p_dir, u0 = lid_bc(cloud, cloud.dim - 1)
    out, iters, gres = [], [], []
    with M.Context(0) as ctx:
        ctx.upload(cloud)
        ctx.weights(3, 3, download=False)
        ctx.fs_create(1.0, 0.0025, 5e-4, 2, 1e-10, 20000, u_dir, u_dir,
                      u_dir if cloud.dim == 3 else None, p_dir, precision=M.MP_PREC_MIXED,
                      inner_tol=1e-5, precond=M.MP_PRECOND_AMG, krylov=1)
        ctx.set_field(M.MP_U, u0)
        for f in (M.MP_V, M.MP_W):
            ctx.set_field(f, np.zeros(cloud.nv))
        ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
        ctx.guess_mode(M.MP_GUESS_PROJECT)
        ctx.guess_window(window)
        with pytest.raises(M.MemphysError):
            ctx.guess_window(25)
        for it in range(steps):
            ctx.step(1, it)
            st = ctx.stats()
            assert st.converged == 1 and st.rel_res <= 1e-10
            iters.append(st.iterations)
            gres.append(st.guess_rel_res)
            out.append([ctx.get_field(f) for f in (M.MP_U, M.MP_V, M.MP_W, M.MP_P)])
    return out, iters, gres


def test_projected_guess_agrees_and_starts_closest(cube3d_small):
    """The projection onto the last 16 solutions reaches the same fields as the
    order-2 extrapolation, and from step 10 on starts from a smaller residual
    and takes fewer iterations than the order-4 extrapolation."""
    steps = 40
    proj, it_p, g_p = _run_project(cube3d_small, steps)
    ext2, it_2 = _run(cube3d_small, steps, 2)
    ext4, it_4 = _run(cube3d_small, steps, 4)
    errs = [rel_err(a, b) for a, b in zip(proj[-1], ext2[-1])]
    _print("projection vs order 2 after", steps, "steps:", errs)
    _print("mean iterations (steps 10-39): projection", np.mean(it_p[10:]), "order 4",
           np.mean(it_4[10:]), "order 2", np.mean(it_2[10:]))
    _print("projection: guess residual, median of steps 10-39", np.median(g_p[10:]))
    assert max(errs) <= 1e-8
    assert np.mean(it_p[10:]) < np.mean(it_4[10:])
