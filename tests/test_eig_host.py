"""CPU checks of the small dense eigen routines behind the GPU Arnoldi
iteration (mp_operator_eigenvalue): eigenvalues of upper Hessenberg matrices
against LAPACK (numpy), eigenvectors by their residual, and a matrix with a
known complex-conjugate spectrum (a rotation-scaling block)."""
import numpy as np
import scipy.linalg as sl


def _match(ev, ref):
    return max(np.min(np.abs(ref - e)) for e in ev) / np.max(np.abs(ref))


def test_hessenberg_eigenvalues_match_lapack():
    import memphys_amd as M
    rng = np.random.default_rng(7)
    worst = 0.0
    for m in (2, 5, 13, 30, 40):
        for _ in range(10):
            H = sl.hessenberg(rng.standard_normal((m, m)))
            ev, y = M.hessenberg_eig(H, which=0)
            ref = np.linalg.eigvals(H)
            worst = max(worst, _match(ev, ref), _match(ref, ev))
            res = np.linalg.norm(H @ y - ev[0] * y) / np.max(np.abs(ev))
            assert abs(np.linalg.norm(y) - 1.0) <= 1e-12
            worst = max(worst, res)
    assert worst <= 1e-12
    # a planted transpose (a different spectrum for non-normal H) fails
    H = sl.hessenberg(rng.standard_normal((12, 12)))
    ev, _ = M.hessenberg_eig(H)
    assert _match(np.linalg.eigvals(H + np.triu(H, 2).T), ev) > 1e-6


def test_known_complex_pair():
    import memphys_amd as M
    # block diagonal [[a, -b], [b, a]] (+) diag(d): eigenvalues a +- ib and d
    a, b = 3.0, 4.0
    H = np.zeros((4, 4))
    H[0, 0], H[0, 1], H[1, 0], H[1, 1] = a, -b, b, a
    H[2, 2], H[3, 3] = -2.0, 1.0
    H[2, 1] = 1e-3  # couples the blocks, stays upper Hessenberg
    ev, _ = M.hessenberg_eig(H)
    ref = np.linalg.eigvals(H)
    assert _match(ev, ref) <= 1e-13
    big = ev[np.argmax(np.abs(ev))]
    assert abs(abs(big) - 5.0) <= 1e-3
