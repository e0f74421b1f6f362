"""GPU parity: the HIP path through the C ABI against the oracle.

Tolerances: stencil weights within 1e-10 relative (norm over each operator);
stencil-operator products within 1e-14 (same arithmetic order as the oracle);
fields after N steps within 1e-8 relative to the FP64 oracle, both sides
solving the pressure equation far below that level.
"""
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


def test_weights_2d_match_oracle(ctx, cavity2d):
    import oracle as O
    _print("case: weights 2D cavity nv", cavity2d.nv)
    ctx.upload(cavity2d)
    (gx, gy, gz, lap), _ = ctx.weights(3, 3)
    ox, oy, _, ol, _ = O.stencil_weights(cavity2d, 3, 3)
    for a, b in ((gx, ox), (gy, oy), (lap, ol)):
        assert rel_err(a, b) <= 1e-10
    exact = np.mean(np.concatenate([gx == ox, gy == oy, lap == ol]))
    _print("bit-exact fraction", exact, "max rel", max(rel_err(gx, ox), rel_err(lap, ol)))


def test_weights_3d_match_oracle(ctx, cube3d_small):
    import oracle as O
    c = cube3d_small
    _print("case: weights 3D cube nv", c.nv, "k", sorted(set(np.diff(c.row_ptr).tolist())))
    ctx.upload(c)
    (gx, gy, gz, lap), _ = ctx.weights(3, 3)
    ox, oy, oz, ol, _ = O.stencil_weights(c, 3, 3)
    for a, b in ((gx, ox), (gy, oy), (gz, oz), (lap, ol)):
        assert rel_err(a, b) <= 1e-10
    # per-stencil relative error (each row's weight vector)
    worst = 0.0
    for r in range(0, c.nv, 37):
        s = slice(c.row_ptr[r], c.row_ptr[r + 1])
        worst = max(worst, rel_err(lap[s], ol[s]), rel_err(gx[s], ox[s]))
    assert worst <= 1e-10
    _print("3D worst per-stencil rel", worst,
           "bit-exact fraction", np.mean(np.concatenate([gx == ox, lap == ol])))


def test_operator_apply_matches_oracle(ctx, cube3d_small):
    import memphys_amd as M
    import oracle as O
    c = cube3d_small
    ctx.upload(c)
    ox, oy, oz, ol, _ = O.stencil_weights(c, 3, 3)
    ctx.set_weights(ox, oy, oz, ol)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(c.nv)
    for op, w in ((M.MP_GX, ox), (M.MP_GY, oy), (M.MP_GZ, oz), (M.MP_LAP, ol)):
        y = ctx.apply(op, x)
        yo = O.spmv(c, w, x)
        assert rel_err(y, yo) <= 1e-14


def _ppe_case(ctx, cloud, precision, krylov=0):
    import memphys_amd as M
    import oracle as O
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    ctx.upload(cloud)
    w = O.stencil_weights(cloud, 3, 3)[:4]
    ctx.set_weights(*w)
    p_dir = np.zeros(cloud.nv, dtype=np.uint8)
    # 1e-10: the 2D cavity's BiCGSTAB tolerance; FP64's attainable true residual on
    # these operators (entries ~1e5) is ~1e-12
    n = ctx.solver_init(p_dir, 0.0, 0.0, 1.0, tol=1e-10, n_iter=20000, precision=precision)
    assert n == cloud.nv + 1
    if krylov:
        ctx.solver_krylov(M.MP_KRYLOV_GMRES)
    ones = np.ones(cloud.nv, dtype=np.uint8)
    fs = O.FractionalStep(cloud, w, ones, ones, ones if cloud.dim == 3 else None, p_dir,
                          1.0, 0.01, 1e-3, 2, solver_kind=0, precond=2, tol=1e-14)
    rp, ci, v = fs.ppe_csr()
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    rng = np.random.default_rng(11)
    b = rng.standard_normal(n)
    b[-1] = 0.0
    x_ref = spl.spsolve(A.tocsc(), b)
    x, xo = ctx.solve(b, np.zeros(n))
    st = ctx.stats()
    _print("PPE", cloud.dim, "D prec", precision, "krylov", krylov, "iters", st.iterations,
           "cycles", st.outer_cycles, "rel_res", st.rel_res, "gmres switches",
           st.gmres_switches)
    if krylov and precision == M.MP_PREC_FP64 and cloud.dim == 2:
        # GMRES(30) with Jacobi stalls on the 2D cavity's operator; the stalled
        # system continues with BiCGSTAB
        assert st.gmres_switches >= 1
    assert st.rel_res <= 1e-10
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-10
    _print("x vs SuperLU rel", rel_err(x, x_ref))
    assert rel_err(x, x_ref) <= 1e-6
    assert xo[-1] == x[-1]


@pytest.mark.parametrize("krylov", [0, 1])
def test_ppe_solve_fp64_matches_superlu(ctx, cavity2d, krylov):
    import memphys_amd as M
    _ppe_case(ctx, cavity2d, M.MP_PREC_FP64, krylov)


@pytest.mark.parametrize("krylov", [0, 1])
def test_ppe_solve_mixed_matches_superlu(ctx, cube3d_small, krylov):
    import memphys_amd as M
    _ppe_case(ctx, cube3d_small, M.MP_PREC_MIXED, krylov)


@pytest.mark.parametrize("prec", [0, 1])
def test_ppe_solve_nonfinite_rhs_is_not_converged(ctx, cavity2d, prec):
    """A NaN in the right-hand side reports a NaN residual and no convergence
    (not a zero residual); the same solver then solves a finite system."""
    import memphys_amd as M
    import oracle as O
    ctx.upload(cavity2d)
    w = O.stencil_weights(cavity2d, 3, 3)[:4]
    ctx.set_weights(*w)
    p_dir = np.zeros(cavity2d.nv, dtype=np.uint8)
    n = ctx.solver_init(p_dir, 0.0, 0.0, 1.0, tol=1e-10, n_iter=2000,
                        precision=M.MP_PREC_MIXED if prec else M.MP_PREC_FP64)
    rng = np.random.default_rng(12)
    b = rng.standard_normal(n)
    b[-1] = 0.0
    bad = b.copy()
    bad[n // 3] = np.nan
    _print("case: NaN right-hand side, precision", prec)
    ctx.solve(bad, np.zeros(n))
    st = ctx.stats()
    _print("iters", st.iterations, "rel_res", st.rel_res, "converged", st.converged)
    assert st.converged == 0 and np.isnan(st.rel_res)
    x, _ = ctx.solve(b, np.zeros(n))
    st = ctx.stats()
    assert st.converged == 1 and st.rel_res <= 1e-10 and np.all(np.isfinite(x))


def _fs_parity(ctx, cloud, steps, lid_axis, rho, mu, dt, precision, precond=0, krylov=0,
               restart=0):
    import memphys_amd as M
    import oracle as O
    ctx.upload(cloud)
    (gx, gy, gz, lap), _ = ctx.weights(3, 3)
    w = (gx, gy, gz, lap)
    u_dir, p_dir, u0 = lid_bc(cloud, lid_axis)
    ctx.fs_create(rho, mu, dt, 2, 1e-12, 20000, u_dir, u_dir, u_dir if cloud.dim == 3 else None,
                  p_dir, precision=precision, precond=precond, krylov=krylov, restart=restart)
    nv = cloud.nv
    ctx.set_field(M.MP_U, u0)
    ctx.set_field(M.MP_V, np.zeros(nv))
    if cloud.dim == 3:
        ctx.set_field(M.MP_W, np.zeros(nv))
    ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
    fs = O.FractionalStep(cloud, w, u_dir, u_dir, u_dir if cloud.dim == 3 else None, p_dir,
                          rho, mu, dt, 2, solver_kind=0, precond=2, tol=1e-14, maxit=20000)
    u, v = u0.copy(), np.zeros(nv)
    ww = np.zeros(nv) if cloud.dim == 3 else None
    p = np.zeros(fs.n_sys)
    for it in range(steps):
        ctx.step(1, it)
        fs.step(it, u, v, ww, p)
    st = ctx.stats()
    errs = [rel_err(ctx.get_field(M.MP_U), u), rel_err(ctx.get_field(M.MP_V), v),
            rel_err(ctx.get_field(M.MP_P), p)]
    if cloud.dim == 3:
        errs.append(rel_err(ctx.get_field(M.MP_W), ww))
    _print(f"FS {cloud.dim}D steps {steps} prec {precision} precond {precond} krylov {krylov}"
           f" restart {restart}: errs {errs}"
           f" iters {st.iterations} total {st.total_iterations} rel_res {st.rel_res}")
    assert max(errs) <= 1e-8
    if krylov and precond == M.MP_PRECOND_AMG:
        assert st.gmres_switches == 0  # no cycle stalls with the multigrid preconditioner
    div = ctx.divergence()
    div_o, _ = O.divergence(cloud, gx, gy, gz, ctx.get_field(M.MP_U), ctx.get_field(M.MP_V),
                            ctx.get_field(M.MP_W) if cloud.dim == 3 else None)
    assert np.allclose(div, div_o, rtol=1e-12, atol=1e-14)
    return st.total_iterations


def test_fs_2d_cavity_matches_oracle(ctx, cavity2d):
    import memphys_amd as M
    # dt = 1e-3 exceeds the explicit stability bound of this scheme on this cloud
    # (nu*dt*|lambda_max(L)| = 3.3); 2e-4 keeps AB2 stable.
    _fs_parity(ctx, cavity2d, 10, 1, 1.0, 0.01, 2e-4, M.MP_PREC_FP64)


def test_fs_3d_cube_matches_oracle(ctx, cube3d_small):
    import memphys_amd as M
    _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, M.MP_PREC_FP64)


def test_fs_3d_cube_mixed_matches_oracle(ctx, cube3d_small):
    import memphys_amd as M
    _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, M.MP_PREC_MIXED)


def test_fs_3d_cube_multigrid_matches_oracle(ctx, cube3d_small):
    """Multigrid-preconditioned BiCGSTAB (FP64 and FP32-inner) reaches the
    same fields as the oracle, in far fewer iterations than Jacobi."""
    import memphys_amd as M
    it_jac = _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, M.MP_PREC_FP64)
    it_mg = _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, M.MP_PREC_FP64,
                       M.MP_PRECOND_AMG)
    info = ctx.precond_info()
    _print("hierarchy", info)
    assert info["levels"] >= 2 and info["rows"][0] == cube3d_small.nv
    it_mx = _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, M.MP_PREC_MIXED,
                       M.MP_PRECOND_AMG)
    _print("iterations jacobi", it_jac, "multigrid fp64", it_mg, "multigrid mixed", it_mx)
    assert it_mg <= 0.5 * it_jac


def test_fs_2d_cavity_multigrid_matches_oracle(ctx, cavity2d):
    import memphys_amd as M
    _fs_parity(ctx, cavity2d, 10, 1, 1.0, 0.01, 2e-4, M.MP_PREC_MIXED, M.MP_PRECOND_AMG)


@pytest.mark.parametrize("precision,precond,restart",
                         [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (1, 1, 4)])
def test_fs_3d_cube_gmres_matches_oracle(ctx, cube3d_small, precision, precond, restart):
    """Restarted GMRES (FP64 or FP32-inner, Jacobi or multigrid, the default
    30 columns or 4, which restarts many times) reaches the oracle's fields."""
    import memphys_amd as M
    _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, precision, precond,
               krylov=M.MP_KRYLOV_GMRES, restart=restart)


def test_gmres_multigrid_needs_fewer_vcycles_than_bicgstab(ctx, cube3d_small):
    """With the multigrid preconditioner a GMRES iteration applies one V-cycle
    and a BiCGSTAB iteration two: GMRES needs fewer V-cycles per solve."""
    import memphys_amd as M
    it_b = _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, M.MP_PREC_MIXED,
                      M.MP_PRECOND_AMG)
    it_g = _fs_parity(ctx, cube3d_small, 8, 2, 1.0, 0.0025, 5e-4, M.MP_PREC_MIXED,
                      M.MP_PRECOND_AMG, krylov=M.MP_KRYLOV_GMRES)
    _print("V-cycles: BiCGSTAB", 2 * it_b, "GMRES", it_g)
    assert it_g < 2 * it_b


def test_fs_2d_cavity_gmres_multigrid_matches_oracle(ctx, cavity2d):
    import memphys_amd as M
    _fs_parity(ctx, cavity2d, 10, 1, 1.0, 0.01, 2e-4, M.MP_PREC_MIXED, M.MP_PRECOND_AMG,
               krylov=M.MP_KRYLOV_GMRES)


@pytest.mark.parametrize("precision,precond", [(0, 0), (1, 1)])
def test_graph_replay_matches_plain_launches(ctx, cube3d_small, precision, precond):
    """Krylov chunks replayed from hipGraphs (many replays per solve, eager
    predictor/corrector work and host polls between them) give bit-identical
    fields to the same chunks launched one by one."""
    import memphys_amd as M
    cloud = cube3d_small
    ctx.upload(cloud)
    ctx.weights(3, 3, download=False)
    u_dir, p_dir, u0 = lid_bc(cloud, 2)
    out = []
    for graphs in (True, False, True):
        ctx.graph_mode(graphs)
        ctx.fs_create(1.0, 0.0025, 5e-4, 2, 1e-10, 20000, u_dir, u_dir, u_dir, p_dir,
                      precision=precision, precond=precond, check_every=4)
        ctx.set_field(M.MP_U, u0)
        for w in (M.MP_V, M.MP_W):
            ctx.set_field(w, np.zeros(cloud.nv))
        ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
        ctx.step(6, 0)
        out.append([ctx.get_field(w) for w in (M.MP_U, M.MP_V, M.MP_W, M.MP_P)])
        _print(f"graphs {graphs} prec {precision} precond {precond}: total iterations",
               ctx.stats().total_iterations)
    ctx.graph_mode(True)
    for a, b, c in zip(*out):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_step_graphs_follow_guess_and_force_changes(ctx, cube3d_small):
    """The fixed parts of a step (predictor, right-hand side, extrapolated
    guess; history, corrector, boundary rows) are replayed from graphs keyed by
    the ping-pong buffers, the Euler/AB2 switch, the guess and the body force:
    a run that switches the guess on, sets a body force and clears it again
    gives the same fields with graphs as with plain launches."""
    import memphys_amd as M
    cloud = cube3d_small
    ctx.upload(cloud)
    ctx.weights(3, 3, download=False)
    u_dir, p_dir, u0 = lid_bc(cloud, 2)
    g = np.sin(3.0 * cloud.xyz[:, 0]) * 0.2
    out = []
    for graphs in (True, False):
        ctx.graph_mode(graphs)
        ctx.fs_create(1.0, 0.0025, 5e-4, 2, 1e-10, 20000, u_dir, u_dir, u_dir, p_dir,
                      precision=M.MP_PREC_MIXED, precond=M.MP_PRECOND_AMG,
                      krylov=M.MP_KRYLOV_GMRES)
        ctx.set_field(M.MP_U, u0)
        for w in (M.MP_V, M.MP_W):
            ctx.set_field(w, np.zeros(cloud.nv))
        ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
        ctx.guess_mode(M.MP_GUESS_PREVIOUS)
        ctx.step(2, 0)
        ctx.guess_mode(M.MP_GUESS_EXTRAPOLATE)
        ctx.step(2, 2)
        ctx.body_force(g, np.zeros(cloud.nv), -g)
        ctx.step(2, 4)
        ctx.body_force()
        ctx.step(2, 6)
        out.append([ctx.get_field(w) for w in (M.MP_U, M.MP_V, M.MP_W, M.MP_P)])
        _print(f"graphs {graphs}: total iterations", ctx.stats().total_iterations)
    ctx.graph_mode(True)
    ctx.guess_mode(M.MP_GUESS_PREVIOUS)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_condition_estimate_matches_oracle(ctx, cube3d_small):
    """cond = ||A||_1 * Hager/Higham estimate of ||A^-1||_1 (the estimator behind
    Eigen's rcond, coefficient_computations.cpp:312), same arithmetic order."""
    import oracle as O
    c = cube3d_small
    ctx.upload(c)
    _, cond = ctx.weights(3, 3, cond=True)
    ocond = O.stencil_weights(c, 3, 3, want_cond=True)[4]
    _print("cond max/min/avg", cond.max(), cond.min(), cond.mean(),
           "bit-exact fraction", np.mean(cond == ocond))
    assert rel_err(cond, ocond) <= 1e-12


CLASS_API_CASES = [
    ("eigen_ilu_bicgstab", None, None),
    ("hypre_ilu_gmres", None, None),
    # the fast path reached from a reference driver through the settings file
    ("hypre_ilu_gmres", "precision,mixed\npreconditioner,multigrid\nguess,extrapolate\n"
                        "inner_tolerance,1e-5\n", None),
    ("hypre_ilu_gmres", "precision,mixed\npreconditioner,multigrid\nguess,extrapolate\n"
                        "extrapolation_order,3\n", None),
    ("hypre_ilu_gmres", "precision,mixed\npreconditioner,multigrid\nguess,project\n"
                        "projection_window,6\n", None),
    # the body-force overload of single_timestep_2d
    ("eigen_ilu_bicgstab", None, 2.0),
    ("hypre_ilu_gmres", "precision,mixed\npreconditioner,multigrid\n", 2.0),
]


@pytest.mark.parametrize("solver_type,settings,gravity", CLASS_API_CASES)
def test_class_api_driver_matches_oracle(tmp_path, solver_type, settings, gravity):
    """examples/lid_driven_cavity (PARAMETERS -> POINTS -> CLOUD ->
    FRACTIONAL_STEP_1 -> time loop) against the oracle on the same cloud, with
    the CSV's solver_type selecting BiCGSTAB or GMRES(gmres_kdim), GPU options
    from a settings file (MEMPHYS_GPU_SETTINGS) and, with `gravity`, the
    body-force overload (fractional_step_1.cpp:65-76)."""
    import os
    import subprocess

    import memphys_amd as M
    import oracle as O
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "examples", "lid_driven_cavity")
    out = str(tmp_path / "fields.bin")
    csv = tmp_path / "parameters.csv"
    text = open(os.path.join(root, "examples", "parameters_cavity.csv")).read()
    csv.write_text(text.replace("solver_type,eigen_ilu_bicgstab", f"solver_type,{solver_type}"))
    env = dict(os.environ)
    env.pop("MEMPHYS_GPU_SETTINGS", None)
    if settings:
        (tmp_path / "gpu.csv").write_text(settings)
        env["MEMPHYS_GPU_SETTINGS"] = str(tmp_path / "gpu.csv")
    steps, dt, Re, n = 10, 1e-3, 100.0, 30
    cmd = [exe, str(csv), "2", str(n), str(steps), str(dt), str(Re), out]
    if gravity is not None:
        cmd.append(str(gravity))
    _print("case: class API", solver_type, repr(settings), "gravity", gravity)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    _print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"solver {solver_type}" in r.stdout
    if settings and "multigrid" in settings:
        assert "precision 1 preconditioner 1" in r.stdout
    raw = open(out, "rb").read()
    nv, dim, nsys = np.frombuffer(raw[:12], dtype=np.int32)
    data = np.frombuffer(raw[12:], dtype=np.float64)
    xyz = data[:nv * dim].reshape(nv, dim)
    u, v, p = data[nv * dim:nv * dim + nv], data[nv * dim + nv:nv * dim + 2 * nv], data[nv * dim + 2 * nv:]
    cloud = M.box_cloud(dim=2, n=n, jitter=0.25, seed=20240601, k_interior=30)
    assert np.array_equal(cloud.xyz, xyz)  # same generator, kNN and RCM order
    w = O.stencil_weights(cloud, 3, 3)[:4]
    u_dir, p_dir, u0 = lid_bc(cloud, 1)
    fs = O.FractionalStep(cloud, w, u_dir, u_dir, None, p_dir, 1.0, 1.0 / Re, dt, 2, precond=2,
                          tol=1e-14, maxit=20000)
    if gravity is not None:
        fs.set_body_force(0.5 * gravity * np.sin(np.pi * cloud.xyz[:, 1]), np.full(nv, -gravity))
    uo, vo, po = u0.copy(), np.zeros(nv), np.zeros(fs.n_sys)
    for it in range(steps):
        fs.step(it, uo, vo, None, po)
    errs = (rel_err(u, uo), rel_err(v, vo), rel_err(p, po))
    _print("class API driver errs", errs)
    assert max(errs) <= 1e-8
    # write_iteration_details (postprocessing_functions.cpp:52-72): one row per step
    import re
    lines = open(out + "_iteration_details.csv").read().splitlines()
    assert lines[0] in ("it,n_iter_actual,rel_res_log,abs_res_log",
                        "it,n_iter_actual,rel_res_log,abs_res_log,regul_alpha_log")
    rows = [row.split(",") for row in lines[1:]]
    assert [int(row[0]) for row in rows] == list(range(steps))
    m = re.search(r"last PPE iterations (\d+) rel_res (\S+)", r.stdout)
    assert int(rows[-1][1]) == int(m.group(1))
    assert float(rows[-1][2]) == pytest.approx(float(m.group(2)), rel=1e-3)
    assert all(float(row[2]) <= 1e-8 for row in rows)


def test_weights_mixed_precision_match_oracle(ctx, cube3d_small, cavity2d):
    """FP32 LU + FP64 refinement (interior stencils) and FP64 (boundary-layer
    stencils): weights within 1e-10 relative of the FP64 oracle."""
    import memphys_amd as M
    import oracle as O
    for c in (cube3d_small, cavity2d):
        ctx.upload(c)
        (gx, gy, gz, lap), _ = ctx.weights(3, 3, precision=M.MP_PREC_MIXED)
        ox, oy, oz, ol, _ = O.stencil_weights(c, 3, 3)
        pairs = [(gx, ox), (gy, oy), (lap, ol)] + ([(gz, oz)] if c.dim == 3 else [])
        worst = 0.0
        for r in range(0, c.nv, 7):
            s = slice(c.row_ptr[r], c.row_ptr[r + 1])
            worst = max(worst, max(rel_err(a[s], b[s]) for a, b in pairs))
        _print(f"mixed weights {c.dim}D worst per-stencil rel {worst:.2e}")
        assert worst <= 1e-10
        for a, b in pairs:
            assert rel_err(a, b) <= 1e-10


def test_gmsh_annulus_matches_reference_sample_output(ctx):
    """GPU weights on the reference's own mesh reproduce the operator
    eigenvalues and RBF condition numbers the reference wrote to its sample
    simulation_details.csv."""
    import memphys_amd as M
    from test_gmsh_known_answers import KA, MESH, _check_eig, _largest_eig
    cloud = M.gmsh_cloud(MESH, KA["cloud_size"])
    ctx.upload(cloud)
    (gx, gy, _, lap), cond = ctx.weights(KA["phs_deg"], KA["poly_deg"], cond=True)
    _print("annulus: nv", cloud.nv, "cond max", cond.max())
    _check_eig(_largest_eig(gx, cloud), KA["grad_x_eigval"], 1e-10)
    _check_eig(_largest_eig(gy, cloud), KA["grad_y_eigval"], 1e-10)
    _check_eig(_largest_eig(lap, cloud), KA["laplace_eigval"], 1e-10)
    assert cond.max() == pytest.approx(KA["rbf_cond_num_max"], rel=1e-12)
    assert cond.min() == pytest.approx(KA["rbf_cond_num_min"], rel=1e-12)
    assert cond.mean() == pytest.approx(KA["rbf_cond_num_avg"], rel=1e-12)
    # the GPU Arnoldi iteration (PARAMETERS::calc_dt's eigenvalues)
    for op, key in ((M.MP_GX, "grad_x_eigval"), (M.MP_GY, "grad_y_eigval"),
                    (M.MP_LAP, "laplace_eigval")):
        ev = ctx.eigenvalue(op)
        _print("GPU Arnoldi", key, ev, "reference", KA[key])
        _check_eig(ev, KA[key], 1e-9)


def test_class_api_gmsh_setup_matches_reference_sample_output(tmp_path):
    """PARAMETERS(csv, msh) -> POINTS -> CLOUD -> calc_dt through the C++ class
    API reproduces the reference's simulation_details.csv for its own mesh."""
    import os
    import subprocess
    from test_gmsh_known_answers import KA, MESH
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "examples", "annulus_operators")
    a = KA["calc_dt_args"]
    golden = os.path.join(root, "tests", "golden")
    sample = os.path.join(golden, "ecc_ratio_0.5_r_i_0.6_n_200_polydeg_3")
    prefix = str(tmp_path / "annulus")
    out = subprocess.run([exe, os.path.join(root, "examples", "parameters_annulus.csv"), MESH,
                          str(a["u0"]), str(a["v0"]), str(a["alpha"]), prefix,
                          sample + "_tecplot.plt"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    got = {}
    for line in out.stdout.splitlines():
        parts = line.split(",")
        if len(parts) >= 2 and parts[0] in KA or parts[0] in ("RBF_cond_num_max",
                                                               "RBF_cond_num_min",
                                                               "RBF_cond_num_avg"):
            got[parts[0]] = [float(x) for x in parts[1:]]
    _print("class API annulus:", got)
    for k in ("nv", "nv_boundary", "nv_internal", "cloud_size"):
        assert got[k][0] == KA[k]
    for k in ("dx_max", "dx_min", "dx_avg"):
        assert got[k][0] == pytest.approx(KA[k], rel=1e-14)
    for k, ref in (("RBF_cond_num_max", "rbf_cond_num_max"), ("RBF_cond_num_min", "rbf_cond_num_min"),
                   ("RBF_cond_num_avg", "rbf_cond_num_avg")):
        assert got[k][0] == pytest.approx(KA[ref], rel=1e-12)
    for k in ("grad_x_eigval", "grad_y_eigval", "laplace_eigval"):
        re, im = got[k]
        assert re == pytest.approx(KA[k][0], rel=1e-9)
        assert abs(im - abs(KA[k][1])) <= 1e-9 * abs(KA[k][0])
    assert got["timestep_dt"][0] == pytest.approx(KA["timestep_dt"], rel=1e-9)
    # read_restart_tecplot_file + write_tecplot_steady_variables: the reference's
    # own tecplot file, read onto the nodes and written back, byte for byte
    # (coordinates of every original vertex, the deleted vertices' values from
    # their nearest node, the triangles)
    mine = open(prefix + "_tecplot.plt", "rb").read()
    ref = open(sample + "_tecplot.plt", "rb").read()
    _print("tecplot round trip:", len(mine), "bytes, identical", mine == ref)
    assert mine == ref
    # write_simulation_details against the reference's file: same lines in the
    # same order; host, user, paths, the step count of its time loop and the
    # timers differ by nature
    mine = open(prefix + "_simulation_details.csv").read().splitlines()
    ref = open(sample + "_simulation_details.csv").read().splitlines()
    mine = [ln for ln in mine if not ln.startswith("restart,file")]
    assert len(mine) == len(ref)
    tol = {"Characteristic": 1e-14, "RBF_cond_num_max": 1e-12, "RBF_cond_num_min": 1e-12,
           "RBF_cond_num_avg": 1e-12, "grad_x_eigval": 1e-9, "grad_y_eigval": 1e-9,
           "laplace_eigval": 1e-9, "timestep_dt": 1e-9}
    skip = ("hostname", "username", "meshfile", "nt_actual")
    timers = ref.index("Run Times,(seconds)")
    for i, (m_, r_) in enumerate(zip(mine, ref)):
        mk, rk = m_.split(","), r_.split(",")
        if i > timers:
            assert mk[0] == rk[0]
            continue
        if rk[0] in skip:
            assert mk[0] == rk[0]
            continue
        key = rk[0]
        if key in tol:
            assert mk[0] == rk[0] and len(mk) == len(rk), (m_, r_)
            nums_m = [float(x) for x in mk[1:] if x not in ("", "seconds")
                      and not x.startswith("dx_")]
            nums_r = [float(x) for x in rk[1:] if x not in ("", "seconds")
                      and not x.startswith("dx_")]
            for x, y in zip(nums_m, nums_r):
                assert x == pytest.approx(y, rel=tol[key], abs=tol[key] * 1e3), (m_, r_)
        else:
            assert m_ == r_, (m_, r_)
