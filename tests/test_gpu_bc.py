"""GPU parity with mixed boundary conditions: a channel with a Dirichlet-
pressure outlet (the non-bordered pressure system of size nv) and a slip wall
whose tangential velocity is Neumann (the corrector's n.grad u = 0 branch).

Reference branches: SOLVER::init's system size and calc_coeff_matrix's
identity rows (solver.cpp:29-38, 172-175), calc_pressure's b = p_old rows
(fractional_step_1.cpp:140-143) and calc_vel_corr's Dirichlet-pressure and
Neumann-velocity rows (fractional_step_1.cpp:159-180).  Every solver form
(FP64 / FP32-inner, Jacobi / multigrid, BiCGSTAB / GMRES) must reach the
oracle's fields to 1e-8 relative after 8 steps, both sides solving to 1e-12
or below.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import memphys_amd as M
    c = M.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def channel2d():
    import memphys_amd as M
    return M.box_cloud(dim=2, n=40, jitter=0.25, seed=11, k_interior=30)


def channel_bc(cloud):
    """Inflow x = 0 (parabolic u, Dirichlet), no-slip bottom (and, in 3D, the
    y faces), slip top (tangential velocity Neumann, normal velocity
    Dirichlet 0, pressure Neumann), outlet x = 1 with Dirichlet pressure
    (p_old there: a linear profile) and no velocity condition."""
    nv, dim = cloud.nv, cloud.dim
    x = cloud.xyz
    b = cloud.boundary == 1
    top_axis = dim - 1
    inlet = b & (np.abs(x[:, 0]) < 1e-12)
    outlet = b & (np.abs(x[:, 0] - 1.0) < 1e-12)
    top = b & (np.abs(x[:, top_axis] - 1.0) < 1e-12)
    u_dir = [np.ones(nv, dtype=np.uint8) for _ in range(dim)]
    for q in range(dim):
        u_dir[q][outlet] = 0
        if q != top_axis:
            u_dir[q][top] = 0  # slip: tangential components Neumann
    p_dir = outlet.astype(np.uint8)
    U = [np.zeros(nv) for _ in range(dim)]
    prof = 4 * x[:, 1] * (1 - x[:, 1])
    if dim == 3:
        prof = prof * 4 * x[:, 2] * (1 - x[:, 2]) + 0.2 * 4 * x[:, 1] * (1 - x[:, 1])
    U[0][inlet] = prof[inlet]
    p0 = np.zeros(nv)
    p0[outlet] = 0.05 * x[outlet, 1]
    return u_dir, p_dir, U, p0


def _run(ctx, cloud, w, steps, precision, precond, krylov, restart=0, rho=1.0, mu=0.01,
         dt=2e-4):
    import memphys_amd as M
    import oracle as O
    dim, nv = cloud.dim, cloud.nv
    u_dir, p_dir, U0, p0 = channel_bc(cloud)
    ctx.fs_create(rho, mu, dt, 2, 1e-12, 20000, u_dir[0], u_dir[1],
                  u_dir[2] if dim == 3 else None, p_dir, precision=precision, precond=precond,
                  krylov=krylov, restart=restart)
    assert ctx.n_sys == nv  # Dirichlet pressure: no border unknown
    for q, f in enumerate((M.MP_U, M.MP_V, M.MP_W)[:dim]):
        ctx.set_field(f, U0[q])
    ctx.set_field(M.MP_P, p0)
    fs = O.FractionalStep(cloud, w, u_dir[0], u_dir[1], u_dir[2] if dim == 3 else None, p_dir,
                          rho, mu, dt, 2, solver_kind=0, precond=2, tol=1e-14, maxit=20000)
    assert fs.n_sys == nv
    U = [u.copy() for u in U0]
    p = p0.copy()
    for it in range(steps):
        ctx.step(1, it)
        fs.step(it, U[0], U[1], U[2] if dim == 3 else None, p)
    st = ctx.stats()
    got = [ctx.get_field(f) for f in (M.MP_U, M.MP_V, M.MP_W)[:dim]]
    pg = ctx.get_field(M.MP_P)
    errs = [rel_err(g, u) for g, u in zip(got, U)] + [rel_err(pg, p)]
    _print(f"channel {dim}D prec {precision} precond {precond} krylov {krylov}: errs "
           f"{['%.2e' % e for e in errs]} iters {st.iterations} total {st.total_iterations} "
           f"rel_res {st.rel_res:.2e} switches {st.gmres_switches}")
    assert max(errs) <= 1e-8
    assert st.converged == 1
    # the outlet keeps its pressure (identity rows, b = p_old: met to the
    # solver's tolerance, as by the reference's Krylov solve) and the slip
    # wall's tangential velocity satisfies n.grad u = 0 (checked through the
    # oracle's equal fields)
    outlet = p_dir == 1
    assert np.abs(pg[outlet] - p0[outlet]).max() <= 1e-10 * np.abs(p0).max()
    return st


@pytest.mark.parametrize("precision,precond,krylov", [
    (0, 0, 0), (0, 0, 1), (1, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 0), (1, 1, 1)])
def test_channel_2d_matches_oracle(ctx, channel2d, precision, precond, krylov):
    import oracle as O
    cloud = channel2d
    ctx.upload(cloud)
    w = O.stencil_weights(cloud, 3, 3)[:4]
    ctx.set_weights(*w)
    _print("case: 2D channel nv", cloud.nv)
    _run(ctx, cloud, w, 8, precision, precond, krylov)


@pytest.mark.parametrize("precision,precond,krylov", [(0, 0, 0), (1, 1, 1), (0, 1, 1)])
def test_channel_3d_matches_oracle(ctx, cube3d_small, precision, precond, krylov):
    import oracle as O
    cloud = cube3d_small
    ctx.upload(cloud)
    w = O.stencil_weights(cloud, 3, 3)[:4]
    ctx.set_weights(*w)
    _print("case: 3D channel nv", cloud.nv)
    _run(ctx, cloud, w, 8, precision, precond, krylov, mu=0.0025, dt=5e-4)


@pytest.mark.parametrize("precision,krylov", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_solver_dirichlet_helmholtz_matches_superlu(ctx, channel2d, precision, krylov):
    """SOLVER::init with Dirichlet rows and unsteady and convection terms
    (solver.cpp:165-214): the GPU solve against SuperLU on the oracle's
    matrix."""
    import memphys_amd as M
    import oracle as O
    import scipy.sparse.linalg as spl
    cloud = channel2d
    ctx.upload(cloud)
    w = O.stencil_weights(cloud, 3, 3)[:4]
    ctx.set_weights(*w)
    d = (cloud.boundary == 1).astype(np.uint8)
    d[np.abs(cloud.xyz[:, 1] - 1.0) < 1e-12] = 0  # Neumann rows on the top face
    coef = (50.0, 0.3, 1.0)
    n = ctx.solver_init(d, *coef, tol=1e-11, n_iter=20000, precision=precision)
    assert n == cloud.nv
    if krylov:
        ctx.solver_krylov(M.MP_KRYLOV_GMRES)
    A = O.coeff_matrix(cloud, w, d, *coef)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n)
    x_ref = spl.spsolve(A.tocsc(), b)
    x, _ = ctx.solve(b, np.zeros(n))
    st = ctx.stats()
    _print(f"Helmholtz prec {precision} krylov {krylov}: iters {st.iterations} "
           f"rel_res {st.rel_res:.2e} x err {rel_err(x, x_ref):.2e}")
    assert st.converged == 1 and st.rel_res <= 1e-11
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-10
    assert rel_err(x, x_ref) <= 1e-7


@pytest.mark.parametrize("precision,precond,krylov", [(0, 0, 0), (1, 1, 1)])
def test_nonfinite_guess_starts_from_zero(ctx, cube3d_small, precision, precond, krylov):
    """A NaN pressure (the solve's guess) with a finite right-hand side: the
    solve starts from zero instead, converges, and the fields stay finite."""
    import memphys_amd as M
    from conftest import lid_bc
    cloud = cube3d_small
    ctx.upload(cloud)
    ctx.weights(3, 3, download=False)
    u_dir, p_dir, u0 = lid_bc(cloud, 2)
    ctx.fs_create(1.0, 0.0025, 5e-4, 2, 1e-10, 20000, u_dir, u_dir, u_dir, p_dir,
                  precision=precision, precond=precond, krylov=krylov)
    ctx.set_field(M.MP_U, u0)
    for f in (M.MP_V, M.MP_W):
        ctx.set_field(f, np.zeros(cloud.nv))
    ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
    ctx.step(2, 0)
    ctx.set_field(M.MP_P, np.full(ctx.n_sys, np.nan))
    _print("case: NaN pressure guess, precision", precision, "precond", precond)
    ctx.step(1, 2)
    st = ctx.stats()
    _print("guess_rel_res", st.guess_rel_res, "rejected", st.guess_rejected, "iters",
           st.iterations, "rel_res", st.rel_res)
    assert np.isnan(st.guess_rel_res) and st.guess_rejected == 1
    assert st.converged == 1 and st.rel_res <= 1e-10
    for f in (M.MP_U, M.MP_V, M.MP_W, M.MP_P):
        assert np.all(np.isfinite(ctx.get_field(f)))
    ctx.step(1, 3)
    assert ctx.stats().guess_rejected == 0 and ctx.stats().converged == 1


@pytest.mark.parametrize("precision,precond,krylov", [(0, 0, 0), (1, 1, 1)])
def test_body_force_matches_oracle(ctx, channel2d, precision, precond, krylov):
    """The body-force overload of single_timestep_2d (fractional_step_1.cpp:
    65-76, 104-118) through mp_fs_body_force: a force that changes from step
    to step, then none, against the oracle's set_body_force; fields to 1e-8."""
    import memphys_amd as M
    import oracle as O
    cloud = channel2d
    ctx.upload(cloud)
    w = O.stencil_weights(cloud, 3, 3)[:4]
    ctx.set_weights(*w)
    nv, x = cloud.nv, cloud.xyz
    u_dir, p_dir, U0, p0 = channel_bc(cloud)
    rho, mu, dt = 1.0, 0.01, 2e-4
    ctx.fs_create(rho, mu, dt, 2, 1e-12, 20000, u_dir[0], u_dir[1], None, p_dir,
                  precision=precision, precond=precond, krylov=krylov)
    for q, f in enumerate((M.MP_U, M.MP_V)):
        ctx.set_field(f, U0[q])
    ctx.set_field(M.MP_P, p0)
    fs = O.FractionalStep(cloud, w, u_dir[0], u_dir[1], None, p_dir, rho, mu, dt, 2,
                          solver_kind=0, precond=2, tol=1e-14, maxit=20000)
    U = [u.copy() for u in U0]
    p = p0.copy()
    _print("case: body force, precision", precision, "precond", precond, "krylov", krylov)
    for it in range(8):
        if it < 6:
            fx = 2.0 * np.sin(np.pi * x[:, 1]) * (1.0 + 0.1 * it)
            fy = -9.81 + 0.5 * x[:, 0]
            ctx.body_force(fx, fy)
            fs.set_body_force(fx, fy)
        else:
            ctx.body_force()
            fs.set_body_force()
        ctx.step(1, it)
        fs.step(it, U[0], U[1], None, p)
    got = [ctx.get_field(f) for f in (M.MP_U, M.MP_V)] + [ctx.get_field(M.MP_P)]
    errs = [rel_err(g, u) for g, u in zip(got, U + [p])]
    st = ctx.stats()
    _print(f"body force errs {['%.2e' % e for e in errs]} rel_res {st.rel_res:.2e}")
    assert max(errs) <= 1e-8 and st.converged == 1
    # the force changed the run: a force-free run differs
    ctx.fs_create(rho, mu, dt, 2, 1e-12, 20000, u_dir[0], u_dir[1], None, p_dir,
                  precision=precision, precond=precond, krylov=krylov)
    for q, f in enumerate((M.MP_U, M.MP_V)):
        ctx.set_field(f, U0[q])
    ctx.set_field(M.MP_P, p0)
    ctx.step(8, 0)
    assert rel_err(ctx.get_field(M.MP_V), U[1]) > 1e-3


def test_body_force_arguments_checked(ctx, channel2d):
    import memphys_amd as M
    cloud = channel2d
    ctx.upload(cloud)
    ctx.weights(3, 3, download=False)
    u_dir, p_dir, U0, p0 = channel_bc(cloud)
    ctx.fs_create(1.0, 0.01, 2e-4, 2, 1e-10, 2000, u_dir[0], u_dir[1], None, p_dir)
    with pytest.raises(M.MemphysError):
        ctx.body_force(np.zeros(cloud.nv))  # fy missing
    with pytest.raises(M.MemphysError):
        ctx.body_force(np.zeros(cloud.nv), np.zeros(cloud.nv), np.zeros(cloud.nv))  # fz in 2D
