"""The projected pressure guess against a least-squares solve on the host.
Each step's guess is the combination of the last K stored solutions whose
residual against the new right-hand side is smallest (P. F. Fischer, Comput.
Methods Appl. Mech. Engrg. 163 (1998) 193-204), so the relative residual of
the guess that the solver reports (guess_rel_res) must equal the minimum
that numpy's least-squares routine finds over the same K solutions, with the
pressure operator from the float64 oracle.  The device forms the Gram matrix
of the time differences of the stored A p on the matrix cores and solves the
column-scaled normal equations on one wavefront; a wrong Gram tile, index or
sign leaves a residual far above the minimum.  Windows 8, 16 and 24 take the
three tilings of the Gram pass (one 16-column block; one block and the
right-hand side alone; two blocks)."""
import numpy as np
import pytest

from conftest import lid_bc

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("window", [8, 16, 24])
def test_projection_matches_host_least_squares(cube3d_small, window):
    import scipy.sparse as sp

    import memphys_amd as M
    import oracle as O
    c = cube3d_small
    steps = window + 8
    u_dir, p_dir, u0 = lid_bc(c, 2)
    press, g_dev = [], []
    with M.Context(0) as ctx:
        ctx.upload(c)
        (gx, gy, gz, lap), _ = ctx.weights(3, 3)
        ctx.fs_create(1.0, 0.0025, 5e-4, 2, 1e-12, 20000, u_dir, u_dir, u_dir, p_dir,
                      precision=M.MP_PREC_MIXED, inner_tol=1e-5, precond=M.MP_PRECOND_AMG,
                      krylov=M.MP_KRYLOV_GMRES)
        ctx.set_field(M.MP_U, u0)
        for f in (M.MP_V, M.MP_W):
            ctx.set_field(f, np.zeros(c.nv))
        ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
        ctx.guess_mode(M.MP_GUESS_PROJECT)
        ctx.guess_window(window)
        n = ctx.n_sys
        for it in range(steps):
            ctx.step(1, it)
            st = ctx.stats()
            assert st.converged == 1 and st.rel_res <= 1e-12
            press.append(ctx.get_field(M.MP_P))
            g_dev.append(st.guess_rel_res)
    fs = O.FractionalStep(c, (gx, gy, gz, lap), u_dir, u_dir, u_dir, p_dir, 1.0, 0.0025, 5e-4, 2)
    rp, ci, v = fs.ppe_csr()
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    ap = [A @ p for p in press]  # A p_s is b_s to the solve's tolerance, 1e-12
    ratios = []
    for s in range(2, steps):
        k = min(window, s)
        # the same span in the basis of backward differences in time, columns
        # scaled to unit norm (the raw solutions are nearly parallel, and an
        # SVD of them drops directions the minimum needs)
        hist = [a.copy() for a in ap[s - k:s][::-1]]  # the newest first
        cols = [hist[0]]
        for o in range(1, k):
            for j in range(k - o):
                hist[j] = hist[j] - hist[j + 1]
            cols.append(hist[0])
        X = np.stack(cols, axis=1)
        X /= np.linalg.norm(X, axis=0)
        b = ap[s]
        beta = np.linalg.lstsq(X, b, rcond=None)[0]
        host = np.linalg.norm(b - X @ beta) / np.linalg.norm(b)
        _print(f"step {s:2d}, {k:2d} solutions: guess residual {g_dev[s]:.3e}, "
               f"least-squares minimum {host:.3e}")
        # the device cannot beat the minimum (beyond b_s - A p_s, 1e-12) and
        # its FP64 normal equations stay within a few percent of it
        assert g_dev[s] >= 0.98 * host - 1e-11
        assert g_dev[s] <= 1.05 * host + 1e-10
        if host > 1e-8:
            ratios.append(g_dev[s] / host)
    assert len(ratios) >= 4
    _print("guess residual / minimum where the minimum is above 1e-8:", np.round(ratios, 4))
    assert np.median(ratios) <= 1.01
