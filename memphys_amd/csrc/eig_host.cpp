// Small dense eigenvalue routines (declarations in eig_host.hpp).
#include "eig_host.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mp_eig {

std::vector<cplx> hessenberg_eigenvalues(const std::vector<double>& Hr, int m) {
    std::vector<cplx> A(Hr.begin(), Hr.end());
    auto at = [&](int r, int c) -> cplx& { return A[(size_t)r * m + c]; };
    std::vector<cplx> ev(m);
    int hi = m - 1, guard = 0;
    while (hi >= 0) {
        if (hi == 0) {
            ev[0] = at(0, 0);
            break;
        }
        // deflation point
        int lo = hi;
        while (lo > 0) {
            const double s = std::abs(at(lo, lo)) + std::abs(at(lo - 1, lo - 1));
            if (std::abs(at(lo, lo - 1)) <= 1e-15 * (s > 0 ? s : 1.0)) {
                at(lo, lo - 1) = 0.0;
                break;
            }
            --lo;
        }
        if (lo == hi) {
            ev[hi] = at(hi, hi);
            --hi;
            guard = 0;
            continue;
        }
        if (++guard > 1000) throw std::runtime_error("eigenvalue iteration did not converge");
        // Wilkinson shift from the trailing 2x2 block
        const cplx a = at(hi - 1, hi - 1), b = at(hi - 1, hi), c = at(hi, hi - 1), d = at(hi, hi);
        const cplx tr = a + d, det = a * d - b * c;
        const cplx disc = std::sqrt(tr * tr - 4.0 * det);
        const cplx l1 = 0.5 * (tr + disc), l2 = 0.5 * (tr - disc);
        cplx mu = (std::abs(l1 - d) < std::abs(l2 - d)) ? l1 : l2;
        if (guard % 11 == 10) mu += 0.5 * std::abs(at(hi, hi - 1));  // exceptional shift
        // one QR step on rows/cols lo..hi with Givens rotations
        std::vector<cplx> cs(hi - lo), sn(hi - lo);
        for (int k = lo; k <= hi; ++k) at(k, k) -= mu;
        for (int k = lo; k < hi; ++k) {
            const cplx x = at(k, k), y = at(k + 1, k);
            const double r = std::sqrt(std::norm(x) + std::norm(y));
            cplx cc = 1.0, ss = 0.0;
            if (r > 0) {
                cc = x / r;
                ss = y / r;
            }
            cs[k - lo] = cc;
            sn[k - lo] = ss;
            for (int j = k; j < m; ++j) {  // G^H applied from the left
                const cplx t1 = at(k, j), t2 = at(k + 1, j);
                at(k, j) = std::conj(cc) * t1 + std::conj(ss) * t2;
                at(k + 1, j) = -ss * t1 + cc * t2;
            }
        }
        for (int k = lo; k < hi; ++k) {  // ... and G from the right
            const cplx cc = cs[k - lo], ss = sn[k - lo];
            for (int i = 0; i <= std::min(hi, k + 2); ++i) {
                const cplx t1 = at(i, k), t2 = at(i, k + 1);
                at(i, k) = t1 * cc + t2 * ss;
                at(i, k + 1) = -t1 * std::conj(ss) + t2 * std::conj(cc);
            }
        }
        for (int k = lo; k <= hi; ++k) at(k, k) += mu;
    }
    return ev;
}

std::vector<cplx> hessenberg_eigenvector(const std::vector<double>& Hr, int m, cplx lam) {
    const cplx shift = lam + cplx(1e-10 * std::max(1.0, std::abs(lam)), 0.0);
    std::vector<cplx> y(m, cplx(1.0, 0.0));
    for (int it = 0; it < 2; ++it) {
        std::vector<cplx> M(Hr.begin(), Hr.end());
        for (int i = 0; i < m; ++i) M[(size_t)i * m + i] -= shift;
        std::vector<cplx> b = y;
        for (int k = 0; k < m; ++k) {
            int p = k;
            for (int i = k + 1; i < m; ++i)
                if (std::abs(M[(size_t)i * m + k]) > std::abs(M[(size_t)p * m + k])) p = i;
            if (p != k) {
                for (int j = 0; j < m; ++j) std::swap(M[(size_t)k * m + j], M[(size_t)p * m + j]);
                std::swap(b[k], b[p]);
            }
            cplx piv = M[(size_t)k * m + k];
            if (std::abs(piv) == 0.0) piv = 1e-300;
            for (int i = k + 1; i < m; ++i) {
                const cplx f = M[(size_t)i * m + k] / piv;
                for (int j = k; j < m; ++j) M[(size_t)i * m + j] -= f * M[(size_t)k * m + j];
                b[i] -= f * b[k];
            }
        }
        for (int k = m - 1; k >= 0; --k) {
            cplx s = b[k];
            for (int j = k + 1; j < m; ++j) s -= M[(size_t)k * m + j] * y[j];
            cplx piv = M[(size_t)k * m + k];
            if (std::abs(piv) == 0.0) piv = 1e-300;
            y[k] = s / piv;
        }
        double nrm = 0.0;
        for (const cplx& v : y) nrm += std::norm(v);
        nrm = std::sqrt(nrm);
        for (cplx& v : y) v /= nrm;
    }
    return y;
}

}  // namespace mp_eig
