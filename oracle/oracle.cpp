// ORACLE — test infrastructure only.  An FP64 CPU restatement of MeMPhyS's
// meshless fractional-step path, used by tests/ (as the checker) and by the
// bench's cpu_baseline leg.  Nothing in the product links or calls this file.
//
// Arithmetic conventions shared with the HIP kernels so that results can be
// compared bit for bit: built with -ffp-contract=off; every multiply-add that
// is meant to fuse is written as fma(); sums run in stencil (CSR) order.
//
// Parity pins (see tests/): numpy/LAPACK golden vectors for the stencil
// weights, polynomial- and RBF-exactness identities, scipy's SuperLU for the
// pressure-Poisson solve, closed-form checks of predictor/corrector.
// The Eigen IncompleteLUT factors and BiCGSTAB iterate sequence of the
// reference cannot be reproduced without Eigen: parity unpinned for those
// intermediates (only the converged solution is compared).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ----------------------------------------------------------------------------
// Graded monomial exponents, parameters.cpp:217-267.  Restated independently of
// the product: degree by degree, every degree-(d-1) term times x, then the
// x-free degree-(d-1) terms times y, then (3D) z^d; in 2D terms times x then y^d.
int or_poly_exponents(int dim, int poly_deg, int* out) {
    if (dim != 2 && dim != 3) return -1;
    std::vector<std::vector<int>> terms;
    terms.push_back(std::vector<int>(dim, 0));
    std::vector<std::vector<int>> prev;
    for (int a = 0; a < dim; ++a) {
        std::vector<int> t(dim, 0);
        t[a] = 1;
        prev.push_back(t);
    }
    if (poly_deg >= 1)
        for (auto& t : prev) terms.push_back(t);
    for (int deg = 2; deg <= poly_deg; ++deg) {
        std::vector<std::vector<int>> cur;
        for (auto t : prev) {
            t[0] += 1;
            cur.push_back(t);
        }
        if (dim == 3)
            for (auto t : prev)
                if (t[0] == 0) {
                    t[1] += 1;
                    cur.push_back(t);
                }
        std::vector<int> last(dim, 0);
        last[dim - 1] = deg;
        cur.push_back(last);
        for (auto& t : cur) terms.push_back(t);
        prev = cur;
    }
    for (size_t r = 0; r < terms.size(); ++r)
        for (int a = 0; a < dim; ++a) out[r * dim + a] = terms[r][a];
    return (int)terms.size();
}

}  // extern "C"

namespace {

// r^q from r^2 for odd integer q (q >= -1 here): sqrt once, then powers of r^2.
inline double rpow(double r2, int q) {
    if (q >= 1) {
        double v = std::sqrt(r2);
        for (int s = 0; s < (q - 1) / 2; ++s) v = v * r2;
        return v;
    }
    double v = 1.0 / std::sqrt(r2);
    for (int s = 0; s < (-q - 1) / 2; ++s) v = v / r2;
    return v;
}

inline double ipow(double x, int e) {
    double v = 1.0;
    for (int s = 0; s < e; ++s) v = v * x;
    return v;
}

// One stencil: assembles A and the gradient/Laplacian right-hand sides of
// coefficient_computations.cpp:33-328 and solves with partial-pivot LU.
// X: k points (AoS, dim), center first.  Outputs k weights per operator.
// Returns the Hager 1-norm estimate of cond(A) when want_cond, else 0.
double stencil_weights(int dim, int phs, int m, const int* exps, int k, const double* X,
                       double* wgx, double* wgy, double* wgz, double* wlap, bool want_cond,
                       int refine) {
    const int M = k + m;
    const int nr = dim + 1;  // rhs columns: gx, gy, (gz), lap
    std::vector<double> xs(k * dim);
    double mn[3], mx[3], sc[3];
    for (int a = 0; a < dim; ++a) mn[a] = mx[a] = X[a];
    for (int j = 0; j < k; ++j)
        for (int a = 0; a < dim; ++a) {
            double v = X[j * dim + a];
            if (mn[a] > v) mn[a] = v;
            if (mx[a] < v) mx[a] = v;
        }
    for (int a = 0; a < dim; ++a) sc[a] = mx[a] - mn[a];
    for (int j = 0; j < k; ++j)
        for (int a = 0; a < dim; ++a) xs[j * dim + a] = (X[j * dim + a] - mn[a]) / sc[a];

    std::vector<double> A(M * M, 0.0), R(M * nr, 0.0);
    auto r2of = [&](int i, int j) {
        double s = 0.0;
        for (int a = 0; a < dim; ++a) {
            double d = xs[i * dim + a] - xs[j * dim + a];
            s = s + d * d;
        }
        return s;
    };
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < k; ++j) {
            double v = rpow(r2of(i, j), phs);
            A[i * M + j] = v;
            A[j * M + i] = v;
        }
    for (int i = 0; i < k; ++i)
        for (int t = 0; t < m; ++t) {
            double v = ipow(xs[i * dim + 0], exps[t * dim + 0]);
            for (int a = 1; a < dim; ++a) v = v * ipow(xs[i * dim + a], exps[t * dim + a]);
            A[i * M + k + t] = v;
            A[(k + t) * M + i] = v;
        }
    const double* xc = &xs[0];
    const double p = (double)phs;
    for (int i = 0; i < k; ++i) {
        double r2 = r2of(i, 0);
        for (int a = 0; a < dim; ++a) {
            double v = (-p) * (xs[i * dim + a] - xc[a]);
            v = v * rpow(r2, phs - 2);
            R[i * nr + a] = v / sc[a];
        }
        double lap = 0.0;
        if (std::fabs(r2) > 1e-14) {
            double g[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < dim; ++a) {
                double d = xs[i * dim + a] - xc[a];
                double d2 = d * d;
                double q4 = rpow(r2, phs - 4);
                double v = (p * p) * d2 * q4;
                v = v - (2.0 * p) * d2 * q4;
                v = v + p * rpow(r2, phs - 2);
                g[a] = v / (sc[a] * sc[a]);
            }
            lap = g[0] + g[1];
            if (dim == 3) lap = lap + g[2];
        }
        R[i * nr + dim] = lap;
    }
    for (int t = 0; t < m; ++t) {
        const int* e = exps + t * dim;
        for (int a = 0; a < dim; ++a) {
            double v = 0.0;
            if (e[a] > 0) {
                if (a == 1) {  // the reference forms d/dy as x^e0 * e1 * y^(e1-1) [* z^e2]
                    v = ipow(xc[0], e[0]);
                    v = v * (double)e[1];
                    v = v * ipow(xc[1], e[1] - 1);
                    if (dim == 3) v = v * ipow(xc[2], e[2]);
                } else if (a == 0) {
                    v = (double)e[0] * ipow(xc[0], e[0] - 1);
                    v = v * ipow(xc[1], e[1]);
                    if (dim == 3) v = v * ipow(xc[2], e[2]);
                } else {
                    v = (double)e[2] * ipow(xc[2], e[2] - 1);
                    v = v * ipow(xc[0], e[0]);
                    v = v * ipow(xc[1], e[1]);
                }
                v = v / sc[a];
            }
            R[(k + t) * nr + a] = v;
        }
        double g[3] = {0.0, 0.0, 0.0};
        if (e[0] > 1) {
            double v = (double)e[0] * ((double)e[0] - 1.0) * ipow(xc[0], e[0] - 2);
            v = v * ipow(xc[1], e[1]);
            if (dim == 3) v = v * ipow(xc[2], e[2]);
            g[0] = v / (sc[0] * sc[0]);
        }
        if (e[1] > 1) {
            double v = (double)e[1] * ((double)e[1] - 1.0) * ipow(xc[1], e[1] - 2);
            v = v * ipow(xc[0], e[0]);
            if (dim == 3) v = v * ipow(xc[2], e[2]);
            g[1] = v / (sc[1] * sc[1]);
        }
        if (dim == 3 && e[2] > 1) {
            double v = (double)e[2] * ((double)e[2] - 1.0) * ipow(xc[2], e[2] - 2);
            v = v * ipow(xc[0], e[0]);
            v = v * ipow(xc[1], e[1]);
            g[2] = v / (sc[2] * sc[2]);
        }
        double lap = g[0] + g[1];
        if (dim == 3) lap = lap + g[2];
        R[(k + t) * nr + dim] = lap;
    }

    std::vector<double> A0;
    std::vector<double> R0;
    if (refine > 0) {
        A0 = A;
        R0 = R;
    }
    double anorm1 = 0.0;
    if (want_cond)
        for (int j = 0; j < M; ++j) {
            double s = 0.0;
            for (int i = 0; i < M; ++i) s += std::fabs(A[i * M + j]);
            anorm1 = std::max(anorm1, s);
        }

    // Right-looking LU with partial pivoting (first maximum on ties).
    std::vector<int> perm(M);
    for (int i = 0; i < M; ++i) perm[i] = i;
    bool singular = false;
    for (int c = 0; c < M; ++c) {
        int piv = c;
        double best = std::fabs(A[c * M + c]);
        for (int i = c + 1; i < M; ++i) {
            double v = std::fabs(A[i * M + c]);
            if (v > best) {
                best = v;
                piv = i;
            }
        }
        if (piv != c) {
            for (int j = 0; j < M; ++j) std::swap(A[c * M + j], A[piv * M + j]);
            for (int j = 0; j < nr; ++j) std::swap(R[c * nr + j], R[piv * nr + j]);
            std::swap(perm[c], perm[piv]);
        }
        double d = A[c * M + c];
        if (d == 0.0) {
            singular = true;
            continue;
        }
        for (int i = c + 1; i < M; ++i) {
            double l = A[i * M + c] / d;
            A[i * M + c] = l;
            for (int j = c + 1; j < M; ++j) A[i * M + j] = std::fma(-l, A[c * M + j], A[i * M + j]);
            for (int j = 0; j < nr; ++j) R[i * nr + j] = std::fma(-l, R[c * nr + j], R[i * nr + j]);
        }
    }
    // back substitution, column oriented
    auto back = [&](std::vector<double>& Rw) {
        for (int c = M - 1; c >= 0; --c)
            for (int j = 0; j < nr; ++j) {
                double x = Rw[c * nr + j] / A[c * M + c];
                Rw[c * nr + j] = x;
                for (int i = 0; i < c; ++i) Rw[i * nr + j] = std::fma(-A[i * M + c], x, Rw[i * nr + j]);
            }
    };
    back(R);
    // optional FP64 iterative refinement with a residual of the unfactored A
    for (int it = 0; it < refine; ++it) {
        std::vector<double> rr(M * nr);
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < nr; ++j) {
                double s = R0[i * nr + j];
                for (int t = 0; t < M; ++t) s = std::fma(-A0[i * M + t], R[t * nr + j], s);
                rr[i * nr + j] = s;
            }
        std::vector<double> rp(M * nr);
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < nr; ++j) rp[i * nr + j] = rr[perm[i] * nr + j];
        for (int c = 0; c < M; ++c)
            for (int i = c + 1; i < M; ++i)
                for (int j = 0; j < nr; ++j)
                    rp[i * nr + j] = std::fma(-A[i * M + c], rp[c * nr + j], rp[i * nr + j]);
        back(rp);
        for (int i = 0; i < M * nr; ++i) R[i] += rp[i];
    }
    for (int j = 0; j < k; ++j) {
        wgx[j] = R[j * nr + 0];
        wgy[j] = R[j * nr + 1];
        if (dim == 3) wgz[j] = R[j * nr + 2];
        wlap[j] = R[j * nr + dim];
    }
    if (!want_cond) return 0.0;
    if (singular) return std::numeric_limits<double>::infinity();
    // Hager (1984) / Higham (1988) estimate of ||A^-1||_1 using the LU factors.
    auto solve = [&](std::vector<double>& v, bool transpose) {
        if (!transpose) {
            std::vector<double> w(M);
            for (int i = 0; i < M; ++i) w[i] = v[perm[i]];
            for (int c = 0; c < M; ++c)
                for (int i = c + 1; i < M; ++i) w[i] = std::fma(-A[i * M + c], w[c], w[i]);
            for (int c = M - 1; c >= 0; --c) {
                w[c] = w[c] / A[c * M + c];
                for (int i = 0; i < c; ++i) w[i] = std::fma(-A[i * M + c], w[c], w[i]);
            }
            v = w;
        } else {  // A^T = U^T L^T P, both triangles column oriented
            std::vector<double> w = v;
            for (int c = 0; c < M; ++c) {
                w[c] = w[c] / A[c * M + c];
                for (int i = c + 1; i < M; ++i) w[i] = std::fma(-A[c * M + i], w[c], w[i]);
            }
            for (int c = M - 1; c >= 0; --c)
                for (int i = 0; i < c; ++i) w[i] = std::fma(-A[c * M + i], w[c], w[i]);
            std::vector<double> o(M);
            for (int i = 0; i < M; ++i) o[perm[i]] = w[i];
            v = o;
        }
    };
    std::vector<double> x(M, 1.0 / M), y, z;
    double est = 0.0;
    int jlast = -1;
    for (int it = 0; it < 5; ++it) {
        y = x;
        solve(y, false);
        double n1 = 0.0;
        for (double t : y) n1 += std::fabs(t);
        est = std::max(est, n1);
        z.assign(M, 0.0);
        for (int i = 0; i < M; ++i) z[i] = (y[i] >= 0.0) ? 1.0 : -1.0;
        solve(z, true);
        int jm = 0;
        double zx = 0.0;
        for (int i = 0; i < M; ++i) {
            if (std::fabs(z[i]) > std::fabs(z[jm])) jm = i;
            zx += z[i] * x[i];
        }
        if (std::fabs(z[jm]) <= zx || jm == jlast) break;
        jlast = jm;
        x.assign(M, 0.0);
        x[jm] = 1.0;
    }
    // Higham's alternating vector
    std::vector<double> alt(M);
    for (int i = 0; i < M; ++i) alt[i] = ((i % 2) ? -1.0 : 1.0) * (1.0 + (double)i / (double)std::max(M - 1, 1));
    solve(alt, false);
    double n1 = 0.0;
    for (double t : alt) n1 += std::fabs(t);
    est = std::max(est, 2.0 * n1 / (3.0 * M));
    return anorm1 * est;
}

}  // namespace

extern "C" {

// Stencil weights for rows `rows[0..nrows)` (all rows when rows == NULL) of a
// CSR stencil set; weights are written at the CSR positions of each row.
// CLOUD::calc_grad_laplace_coeffs, cloud.cpp:207-279.
int or_stencil_weights(int dim, int phs_deg, int m, const int* exps, const double* xyz,
                       const int* row_ptr, const int* col, int nrows, const int* rows, int nv,
                       double* gx, double* gy, double* gz, double* lap, double* cond,
                       int refine, int n_threads) {
    if (dim != 2 && dim != 3) return -1;
    int count = rows ? nrows : nv;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int t = 0; t < count; ++t) {
        int r = rows ? rows[t] : t;
        int b = row_ptr[r], k = row_ptr[r + 1] - row_ptr[r];
        std::vector<double> X(k * dim);
        for (int j = 0; j < k; ++j)
            for (int a = 0; a < dim; ++a) X[j * dim + a] = xyz[col[b + j] * dim + a];
        double cn = stencil_weights(dim, phs_deg, m, exps, k, X.data(), gx + b, gy + b,
                                    dim == 3 ? gz + b : nullptr, lap + b, cond != nullptr, refine);
        if (cond) cond[rows ? t : r] = cn;
    }
    return 0;
}

// Single-stencil entry (points given directly, center first): used by the
// golden-vector tests.  Outputs k weights per operator.
double or_single_stencil(int dim, int phs_deg, int m, const int* exps, int k, const double* X,
                         double* gx, double* gy, double* gz, double* lap, int refine) {
    return stencil_weights(dim, phs_deg, m, exps, k, X, gx, gy, gz, lap, true, refine);
}

// y = Op x over the CSR stencil operator (entry order = stencil order, fma chain).
void or_spmv(int nv, const int* row_ptr, const int* col, const double* val, const double* x,
             double* y) {
    for (int i = 0; i < nv; ++i) {
        double s = 0.0;
        for (int t = row_ptr[i]; t < row_ptr[i + 1]; ++t) s = std::fma(val[t], x[col[t]], s);
        y[i] = s;
    }
}

}  // extern "C"

// ----------------------------------------------------------------------------
// Pressure-Poisson operator, BiCGSTAB and ILUT.
namespace {

struct Csr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

// SOLVER::calc_coeff_matrix, solver.cpp:165-214.  Rows in stencil order; in
// the full-Neumann case each of the first nv rows gets a trailing 1 in column
// nv, and row nv holds nv ones followed by the 0 corner.
Csr build_coeff(int nv, int dim, const int* rp, const int* ci, const double* gx, const double* gy,
                const double* gz, const double* lap, const double* normal, const uint8_t* bnd,
                const uint8_t* dir, double unsteady, double conv, double diff, bool bordered) {
    Csr A;
    A.n = bordered ? nv + 1 : nv;
    A.rp.push_back(0);
    for (int i = 0; i < nv; ++i) {
        if (bnd[i]) {
            if (dir[i]) {
                A.ci.push_back(i);
                A.v.push_back(1.0);
            } else {
                for (int t = rp[i]; t < rp[i + 1]; ++t) {
                    double val = normal[dim * i] * gx[t] + normal[dim * i + 1] * gy[t];
                    if (dim == 3) val += normal[dim * i + 2] * gz[t];
                    A.ci.push_back(ci[t]);
                    A.v.push_back(val);
                }
            }
        } else {
            for (int t = rp[i]; t < rp[i + 1]; ++t) {
                double val = diff * lap[t];
                val += conv * gx[t];
                val += conv * gy[t];
                if (dim == 3) val += conv * gz[t];
                if (ci[t] == i) val += unsteady;
                A.ci.push_back(ci[t]);
                A.v.push_back(val);
            }
        }
        if (bordered) {
            A.ci.push_back(nv);
            A.v.push_back(1.0);
        }
        A.rp.push_back((int)A.ci.size());
    }
    if (bordered) {
        for (int i = 0; i < nv; ++i) {
            A.ci.push_back(i);
            A.v.push_back(1.0);
        }
        A.ci.push_back(nv);
        A.v.push_back(0.0);
        A.rp.push_back((int)A.ci.size());
    }
    return A;
}

// Row loops run in parallel (OpenMP, or_set_threads); each row keeps its own
// sequential fma order, so results do not depend on the thread count.
void spmv(const Csr& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A.n; ++i) {
        double s = 0.0;
        for (int t = A.rp[i]; t < A.rp[i + 1]; ++t) s = std::fma(A.v[t], x[A.ci[t]], s);
        y[i] = s;
    }
}

// Dot product over 256 fixed contiguous chunks (each summed in order, then
// the chunk sums in order): the same value for any thread count.
double dot(const std::vector<double>& a, const std::vector<double>& b) {
    constexpr int NCH = 256;
    const long n = (long)a.size();
    double part[NCH];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCH; ++c) {
        const long i0 = n * c / NCH, i1 = n * (c + 1) / NCH;
        double s = 0.0;
        for (long i = i0; i < i1; ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < NCH; ++c) s += part[c];
    return s;
}

// ILUT(droptol, fill) after Saad, "ILUT: a dual threshold incomplete LU
// factorization" (1994), with the rules of Eigen 3.3.7's IncompleteLUT (the
// reference's preconditioner, solver.cpp:66, 156): a multiplier is dropped when
// |l_ik| <= droptol; an upper entry when |u_ij| <= droptol * |a_i|_2; each row
// keeps at most fill_in / 2 multipliers and fill_in / 2 upper entries besides
// the diagonal (the largest), fill_in = nnz(A) * fill / n + 1 (at most n); a
// zero diagonal becomes sqrt(droptol) * |a_i|_2.  Eigen first orders the
// unknowns by AMD on A + A^T; this restatement keeps the stencil set's (RCM)
// order, so its factors are not Eigen's (parity unpinned).
struct Ilut {
    int n = 0;
    std::vector<int> lrp, lci, urp, uci;  // L strictly lower (unit diag), U with diagonal first
    std::vector<double> lv, uv;

    void compute(const Csr& A, double droptol, double fillfactor) {
        n = A.n;
        long nnz = (long)A.rp[n];
        long fill_in = (long)((double)nnz * fillfactor / (double)n) + 1;
        if (fill_in > n) fill_in = n;
        const int fill = (int)(fill_in / 2);
        lrp.assign(1, 0);
        urp.assign(1, 0);
        lci.clear(); lv.clear(); uci.clear(); uv.clear();
        std::vector<double> w(n, 0.0);
        std::vector<char> in(n, 0);
        std::vector<int> nzl, nzu;
        std::vector<int> udiag_pos;  // position of U(i,i) (first entry of U row)
        for (int i = 0; i < n; ++i) {
            double rn = 0.0;
            nzl.clear(); nzu.clear();
            for (int t = A.rp[i]; t < A.rp[i + 1]; ++t) {
                int c = A.ci[t];
                w[c] += A.v[t];
                rn += A.v[t] * A.v[t];
                if (!in[c]) {
                    in[c] = 1;
                    if (c < i) nzl.push_back(c); else nzu.push_back(c);
                }
            }
            rn = std::sqrt(rn);
            if (rn == 0.0) rn = 1.0;
            // eliminate in increasing column order (nzl grows as fill appears)
            std::make_heap(nzl.begin(), nzl.end(), std::greater<int>());
            std::vector<int> lkeep;
            while (!nzl.empty()) {
                std::pop_heap(nzl.begin(), nzl.end(), std::greater<int>());
                int kcol = nzl.back();
                nzl.pop_back();
                double lik = w[kcol] / uv[urp[kcol]];
                if (std::fabs(lik) <= droptol) {
                    w[kcol] = 0.0;
                    in[kcol] = 0;
                    continue;
                }
                w[kcol] = lik;
                lkeep.push_back(kcol);
                for (int t = urp[kcol] + 1; t < urp[kcol + 1]; ++t) {
                    int c = uci[t];
                    w[c] = std::fma(-lik, uv[t], w[c]);
                    if (!in[c]) {
                        in[c] = 1;
                        if (c < i) {
                            nzl.push_back(c);
                            std::push_heap(nzl.begin(), nzl.end(), std::greater<int>());
                        } else {
                            nzu.push_back(c);
                        }
                    }
                }
            }
            // L part: keep the `fill` largest
            if ((int)lkeep.size() > fill) {
                std::nth_element(lkeep.begin(), lkeep.begin() + fill, lkeep.end(),
                                 [&](int a, int b) { return std::fabs(w[a]) > std::fabs(w[b]); });
                for (size_t t = fill; t < lkeep.size(); ++t) { w[lkeep[t]] = 0.0; in[lkeep[t]] = 0; }
                lkeep.resize(fill);
            }
            std::sort(lkeep.begin(), lkeep.end());
            for (int c : lkeep) { lci.push_back(c); lv.push_back(w[c]); w[c] = 0.0; in[c] = 0; }
            lrp.push_back((int)lci.size());
            // U part: diagonal always kept, off-diagonals dropped by threshold then by count
            double diag = in[i] ? w[i] : 0.0;
            std::vector<int> ukeep;
            for (int c : nzu) {
                if (c == i) continue;
                if (std::fabs(w[c]) > droptol * rn) ukeep.push_back(c);
                else { w[c] = 0.0; in[c] = 0; }
            }
            if ((int)ukeep.size() > fill) {
                std::nth_element(ukeep.begin(), ukeep.begin() + fill, ukeep.end(),
                                 [&](int a, int b) { return std::fabs(w[a]) > std::fabs(w[b]); });
                for (size_t t = fill; t < ukeep.size(); ++t) { w[ukeep[t]] = 0.0; in[ukeep[t]] = 0; }
                ukeep.resize(fill);
            }
            std::sort(ukeep.begin(), ukeep.end());
            if (diag == 0.0) diag = std::sqrt(droptol) * rn;
            uci.push_back(i);
            uv.push_back(diag);
            w[i] = 0.0; in[i] = 0;
            for (int c : ukeep) { uci.push_back(c); uv.push_back(w[c]); w[c] = 0.0; in[c] = 0; }
            urp.push_back((int)uci.size());
        }
    }
    void solve(const double* b, double* x) const {
        for (int i = 0; i < n; ++i) {
            double s = b[i];
            for (int t = lrp[i]; t < lrp[i + 1]; ++t) s = std::fma(-lv[t], x[lci[t]], s);
            x[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = x[i];
            for (int t = urp[i] + 1; t < urp[i + 1]; ++t) s = std::fma(-uv[t], x[uci[t]], s);
            x[i] = s / uv[urp[i]];
        }
    }
};

struct Precond {
    int kind = 0;  // 0 none, 1 Jacobi, 2 ILUT
    std::vector<double> dinv;
    Ilut ilut;
    void setup(const Csr& A, int k, double droptol, double fill) {
        kind = k;
        if (k == 1) {
            dinv.assign(A.n, 1.0);
            for (int i = 0; i < A.n; ++i)
                for (int t = A.rp[i]; t < A.rp[i + 1]; ++t)
                    if (A.ci[t] == i && A.v[t] != 0.0) dinv[i] = 1.0 / A.v[t];
        } else if (k == 2) {
            ilut.compute(A, droptol, fill);
        }
    }
    void apply(const std::vector<double>& in, std::vector<double>& out) const {
        out.resize(in.size());
        if (kind == 0) out = in;
        else if (kind == 1) {
            const long n = (long)in.size();
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) out[i] = dinv[i] * in[i];
        }
        else ilut.solve(in.data(), out.data());
    }
};

// BiCGSTAB as in Eigen 3.3.7 (BiCGSTAB.h): right-preconditioned, recursive
// residual, restart when r0.r underflows relative to |r0|^2, stop when
// |r|^2 <= tol^2 |b|^2 or after maxit iterations.
int bicgstab(const Csr& A, const Precond& P, const std::vector<double>& b, std::vector<double>& x,
             double tol, int maxit, double* err_out) {
    const int n = A.n;
    std::vector<double> r(n), r0, v(n, 0.0), p(n, 0.0), y, z, s(n), t(n), tmp(n);
    spmv(A, x.data(), tmp.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
    r0 = r;
    double r0_sq = dot(r0, r0), rhs_sq = dot(b, b);
    if (rhs_sq == 0.0) {
        std::fill(x.begin(), x.end(), 0.0);
        if (err_out) *err_out = 0.0;
        return 0;
    }
    double rho = 1, alpha = 1, w = 1;
    const double eps = std::numeric_limits<double>::epsilon();
    double tol2 = tol * tol * rhs_sq, eps2 = eps * eps;
    int i = 0, restarts = 0;
    while (dot(r, r) > tol2 && i < maxit) {
        double rho_old = rho;
        rho = dot(r0, r);
        if (std::fabs(rho) < eps2 * r0_sq) {
            spmv(A, x.data(), tmp.data());
            for (int q = 0; q < n; ++q) r[q] = b[q] - tmp[q];
            r0 = r;
            rho = r0_sq = dot(r, r);
            if (restarts++ == 0) i = 0;
        }
        double beta = (rho / rho_old) * (alpha / w);
#pragma omp parallel for schedule(static)
        for (int q = 0; q < n; ++q) p[q] = r[q] + beta * (p[q] - w * v[q]);
        P.apply(p, y);
        spmv(A, y.data(), v.data());
        alpha = rho / dot(r0, v);
#pragma omp parallel for schedule(static)
        for (int q = 0; q < n; ++q) s[q] = r[q] - alpha * v[q];
        P.apply(s, z);
        spmv(A, z.data(), t.data());
        double tt = dot(t, t);
        w = (tt > 0.0) ? dot(t, s) / tt : 0.0;
#pragma omp parallel for schedule(static)
        for (int q = 0; q < n; ++q) {
            x[q] += alpha * y[q] + w * z[q];
            r[q] = s[q] - w * t[q];
        }
        ++i;
    }
    if (err_out) *err_out = std::sqrt(dot(r, r) / rhs_sq);
    return i;
}

// Dense LU solve of the assembled operator (a stand-in for the reference's
// eigen_direct SparseLU branch, solver.cpp:55-60,323-336, for small systems).
void dense_solve(const Csr& A, const std::vector<double>& b, std::vector<double>& x) {
    const int n = A.n;
    std::vector<double> D((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int t = A.rp[i]; t < A.rp[i + 1]; ++t) D[(size_t)i * n + A.ci[t]] += A.v[t];
    x = b;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int i = c + 1; i < n; ++i)
            if (std::fabs(D[(size_t)i * n + c]) > std::fabs(D[(size_t)piv * n + c])) piv = i;
        if (piv != c) {
            for (int j = 0; j < n; ++j) std::swap(D[(size_t)c * n + j], D[(size_t)piv * n + j]);
            std::swap(x[c], x[piv]);
        }
        double d = D[(size_t)c * n + c];
        for (int i = c + 1; i < n; ++i) {
            double l = D[(size_t)i * n + c] / d;
            if (l == 0.0) continue;
            for (int j = c + 1; j < n; ++j) D[(size_t)i * n + j] -= l * D[(size_t)c * n + j];
            x[i] -= l * x[c];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        double s = x[c];
        for (int j = c + 1; j < n; ++j) s -= D[(size_t)c * n + j] * x[j];
        x[c] = s / D[(size_t)c * n + c];
    }
}

}  // namespace

// ----------------------------------------------------------------------------
// FRACTIONAL_STEP_1 (fractional_step_1.cpp:22-222), generalised to 3D by adding
// w, Gz and n_z in the pattern of the x/y terms.
struct OrFs {
    int nv, dim, order;
    double rho, mu, dt;
    std::vector<int> rp, ci;
    std::vector<double> gx, gy, gz, lap, normal;
    std::vector<uint8_t> bnd, udir[3], pdir;
    bool full_neumann;
    Csr A;
    Precond P;
    int solver_kind;  // 0 BiCGSTAB, 1 dense direct
    double tol;
    int maxit;
    std::vector<double> uh[3], S_oo[3];
    std::vector<double> force[3];  // body force (empty: none)
    // logs of the last step
    int last_iters = 0;
    double last_rel = 0, last_abs = 0;
};

namespace {
void op_apply(const OrFs& F, const std::vector<double>& w, const double* x, double* y,
              bool interior_only) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < F.nv; ++i) {
        if (interior_only && F.bnd[i]) { y[i] = 0.0; continue; }
        double s = 0.0;
        for (int t = F.rp[i]; t < F.rp[i + 1]; ++t) s = std::fma(w[t], x[F.ci[t]], s);
        y[i] = s;
    }
}
}  // namespace

extern "C" {

// Threads of the OpenMP loops (time step and weights); results do not depend on it.
void or_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

// solver_kind: 0 = BiCGSTAB, 1 = dense LU.  precond: 0 none, 1 Jacobi, 2 ILUT.
OrFs* or_fs_create(int nv, int dim, const int* rp, const int* ci, const double* gx,
                   const double* gy, const double* gz, const double* lap, const double* normal,
                   const uint8_t* bnd, const uint8_t* udir, const uint8_t* vdir,
                   const uint8_t* wdir, const uint8_t* pdir, double rho, double mu, double dt,
                   int temporal_order, int solver_kind, int precond, double tol, int maxit,
                   double droptol, double fill) {
    OrFs* F = new OrFs();
    F->nv = nv; F->dim = dim; F->order = temporal_order;
    F->rho = rho; F->mu = mu; F->dt = dt;
    int W = rp[nv];
    F->rp.assign(rp, rp + nv + 1);
    F->ci.assign(ci, ci + W);
    F->gx.assign(gx, gx + W);
    F->gy.assign(gy, gy + W);
    if (dim == 3) F->gz.assign(gz, gz + W); else F->gz.assign(W, 0.0);
    F->lap.assign(lap, lap + W);
    F->normal.assign(normal, normal + nv * dim);
    F->bnd.assign(bnd, bnd + nv);
    F->udir[0].assign(udir, udir + nv);
    F->udir[1].assign(vdir, vdir + nv);
    if (dim == 3) F->udir[2].assign(wdir, wdir + nv);
    F->pdir.assign(pdir, pdir + nv);
    F->full_neumann = true;
    for (int i = 0; i < nv; ++i)
        if (bnd[i] && pdir[i]) { F->full_neumann = false; break; }
    F->A = build_coeff(nv, dim, rp, ci, gx, gy, dim == 3 ? gz : nullptr, lap, normal, bnd, pdir,
                       0.0, 0.0, 1.0, F->full_neumann);
    F->solver_kind = solver_kind;
    F->tol = tol;
    F->maxit = maxit;
    if (solver_kind == 0) F->P.setup(F->A, precond, droptol, fill);
    for (int c = 0; c < dim; ++c) {
        F->uh[c].assign(nv, 0.0);
        F->S_oo[c].assign(nv, 0.0);
    }
    return F;
}

void or_fs_destroy(OrFs* F) { delete F; }

// Body force of the following steps (single_timestep_2d's body_force_x/y
// overload, fractional_step_1.cpp:65-76; fz in 3D); fx == NULL removes it.
void or_fs_set_body_force(OrFs* F, const double* fx, const double* fy, const double* fz) {
    const double* f[3] = {fx, fy, fz};
    for (int c = 0; c < 3; ++c) {
        if (fx && c < F->dim && f[c]) F->force[c].assign(f[c], f[c] + F->nv);
        else F->force[c].clear();
    }
}

int or_fs_system_size(OrFs* F) { return F->A.n; }

// Export the assembled PPE operator (CSR) for library cross-checks.
int or_fs_ppe_nnz(OrFs* F) { return F->A.rp[F->A.n]; }
void or_fs_ppe_export(OrFs* F, int* rp, int* ci, double* v) {
    std::copy(F->A.rp.begin(), F->A.rp.end(), rp);
    std::copy(F->A.ci.begin(), F->A.ci.end(), ci);
    std::copy(F->A.v.begin(), F->A.v.end(), v);
}

// SOLVER::init's system (solver.cpp:29-38) and calc_coeff_matrix (165-214)
// for any (unsteady, conv, diff): bordered to nv + 1 unless a boundary node
// is Dirichlet.  or_coeff_export with null outputs returns the entry count;
// otherwise fills rp (n + 1), ci, v and returns n.
int or_coeff_export(int nv, int dim, const int* rp, const int* ci, const double* gx,
                    const double* gy, const double* gz, const double* lap, const double* normal,
                    const uint8_t* bnd, const uint8_t* dir, double unsteady, double conv,
                    double diff, int* out_rp, int* out_ci, double* out_v) {
    bool bordered = true;
    for (int i = 0; i < nv; ++i)
        if (bnd[i] && dir[i]) { bordered = false; break; }
    Csr A = build_coeff(nv, dim, rp, ci, gx, gy, dim == 3 ? gz : nullptr, lap, normal, bnd, dir,
                        unsteady, conv, diff, bordered);
    if (!out_rp || !out_ci || !out_v) return A.rp[A.n];
    std::copy(A.rp.begin(), A.rp.end(), out_rp);
    std::copy(A.ci.begin(), A.ci.end(), out_ci);
    std::copy(A.v.begin(), A.v.end(), out_v);
    return A.n;
}

// SOLVER::general_solve (solver.cpp:283-363), eigen_ilu_bicgstab / direct
// branches: x_new = solve(rhs) from guess x_old; x_old[nv] = x_new[nv].
int or_ppe_solve(OrFs* F, const double* rhs, double* x_new, double* x_old, double* rel_res,
                 double* abs_res) {
    const int n = F->A.n;
    std::vector<double> b(rhs, rhs + n), x(x_old, x_old + n);
    int iters = 1;
    if (F->solver_kind == 0) {
        double err = 0;
        iters = bicgstab(F->A, F->P, b, x, F->tol, F->maxit, &err);
    } else {
        dense_solve(F->A, b, x);
    }
    std::vector<double> ax(n);
    spmv(F->A, x.data(), ax.data());
    double an = 0.0, bn = 0.0;
    for (int i = 0; i < n; ++i) {
        an += (ax[i] - b[i]) * (ax[i] - b[i]);
        bn += b[i] * b[i];
    }
    an = std::sqrt(an);
    bn = std::sqrt(bn);
    if (abs_res) *abs_res = an;
    if (rel_res) *rel_res = an / bn;
    for (int i = 0; i < n; ++i) x_new[i] = x[i];
    if (F->full_neumann) x_old[F->nv] = x_new[F->nv];
    F->last_iters = iters;
    F->last_rel = rel_res ? *rel_res : 0;
    F->last_abs = an;
    return iters;
}

}  // extern "C"

namespace {

// calc_vel_hat (fractional_step_1.cpp:120-134) and calc_pressure's right-hand
// side (fractional_step_1.cpp:136-149).  Writes F->uh, S (dim x nv), b.
void fs_predict_rhs(OrFs* F, int it, double* const U[3], const double* p,
                    std::vector<double> S[3], std::vector<double>& b) {
    const int nv = F->nv, dim = F->dim;
    const double rho = F->rho, mu = F->mu, dt = F->dt;
    const double *u = U[0], *v = U[1], *w = U[2];
    std::vector<double> g[4];
    for (auto& t : g) t.resize(nv);
    for (int c = 0; c < dim; ++c) {
        op_apply(*F, F->gx, U[c], g[0].data(), false);
        op_apply(*F, F->gy, U[c], g[1].data(), false);
        if (dim == 3) op_apply(*F, F->gz, U[c], g[2].data(), false);
        op_apply(*F, F->lap, U[c], g[3].data(), false);
        S[c].resize(nv);
        for (int i = 0; i < nv; ++i) {
            double conv = -(u[i] * g[0][i]) - v[i] * g[1][i];
            if (dim == 3) conv = conv - w[i] * g[2][i];
            S[c][i] = rho * conv + mu * g[3][i];
            // the body-force overload (fractional_step_1.cpp:104-118): + f
            if (!F->force[c].empty()) S[c][i] = S[c][i] + F->force[c][i];
        }
        for (int i = 0; i < nv; ++i) {
            if (F->order == 1 || it == 0)
                F->uh[c][i] = U[c][i] + (dt / rho) * S[c][i];
            else
                F->uh[c][i] = U[c][i] + (dt / rho) * (1.5 * S[c][i] - 0.5 * F->S_oo[c][i]);
        }
    }
    const int n = F->A.n;
    b.assign(n, 0.0);
    std::vector<double> gi(nv);
    op_apply(*F, F->gx, F->uh[0].data(), b.data(), true);
    op_apply(*F, F->gy, F->uh[1].data(), gi.data(), true);
    for (int i = 0; i < nv; ++i) b[i] = b[i] + gi[i];
    if (dim == 3) {
        op_apply(*F, F->gz, F->uh[2].data(), gi.data(), true);
        for (int i = 0; i < nv; ++i) b[i] = b[i] + gi[i];
    }
    for (int i = 0; i < nv; ++i) b[i] = b[i] * rho / dt;
    for (int i = 0; i < nv; ++i)
        if (F->bnd[i]) {
            if (F->pdir[i]) b[i] = p[i];
            else {
                double s = (U[0][i] - F->uh[0][i]) * F->normal[dim * i] +
                           (U[1][i] - F->uh[1][i]) * F->normal[dim * i + 1];
                if (dim == 3) s = s + (U[2][i] - F->uh[2][i]) * F->normal[dim * i + 2];
                b[i] = -rho * s / dt;
            }
        }
    if (n == nv + 1) b[nv] = 0.0;
}

// calc_vel_corr (fractional_step_1.cpp:153-209): un = uh - dt/rho G p, then
// the boundary rows with Neumann pressure.
void fs_correct(OrFs* F, double* const U[3], const double* pn, std::vector<double> un[3]) {
    const int nv = F->nv, dim = F->dim;
    const double rho = F->rho, dt = F->dt;
    std::vector<double> g(nv);
    const std::vector<double>* G[3] = {&F->gx, &F->gy, &F->gz};
    for (int c = 0; c < dim; ++c) {
        op_apply(*F, *G[c], pn, g.data(), false);
        un[c].resize(nv);
        for (int i = 0; i < nv; ++i) un[c][i] = F->uh[c][i] - (dt * g[i]) / rho;
    }
    for (int c = 0; c < dim; ++c)
        for (int i = 0; i < nv; ++i)
            if (F->bnd[i] && !F->pdir[i]) {
                if (F->udir[c][i]) un[c][i] = U[c][i];
                else {
                    double rhs = 0.0, diag = 1.0;
                    for (int t = F->rp[i]; t < F->rp[i + 1]; ++t) {
                        double cf = F->gx[t] * F->normal[dim * i] + F->gy[t] * F->normal[dim * i + 1];
                        if (dim == 3) cf = cf + F->gz[t] * F->normal[dim * i + 2];
                        if (F->ci[t] == i) diag = cf;
                        else rhs = rhs - cf * un[c][F->ci[t]];
                    }
                    un[c][i] = rhs / diag;
                }
            }
}

}  // namespace

extern "C" {

// One FRACTIONAL_STEP_1::single_timestep_2d (3D generalisation when dim == 3)
// followed by the driver's old <- new swap (batch_ml.cpp:91).  Fields u, v, w
// (w ignored in 2D) of size nv and p of size system_size hold the *old*
// values on entry and the new ones on return.
int or_fs_step(OrFs* F, int it, double* u, double* v, double* w, double* p) {
    const int dim = F->dim;
    double* U[3] = {u, v, w};
    std::vector<double> S[3], b;
    fs_predict_rhs(F, it, U, p, S, b);
    std::vector<double> pn(F->A.n);
    double rr, ra;
    or_ppe_solve(F, b.data(), pn.data(), p, &rr, &ra);
    std::vector<double> un[3];
    fs_correct(F, U, pn.data(), un);
    // extras (fractional_step_1.cpp:211-222): AB2 history
    for (int c = 0; c < dim; ++c) F->S_oo[c] = S[c];
    for (int c = 0; c < dim; ++c) std::copy(un[c].begin(), un[c].end(), U[c]);
    std::copy(pn.begin(), pn.end(), p);
    return F->last_iters;
}

// Predictor + pressure right-hand side only (no solve, no state change):
// uh_out dim*nv (component-major), b_out system size.
void or_fs_predict(OrFs* F, int it, double* u, double* v, double* w, const double* p,
                   double* uh_out, double* b_out) {
    double* U[3] = {u, v, w};
    std::vector<double> S[3], b;
    fs_predict_rhs(F, it, U, p, S, b);
    for (int c = 0; c < F->dim; ++c) std::copy(F->uh[c].begin(), F->uh[c].end(), uh_out + (size_t)c * F->nv);
    std::copy(b.begin(), b.end(), b_out);
}

// Corrector given uh (dim*nv, component-major), old fields and a pressure:
// un_out dim*nv.
void or_fs_correct(OrFs* F, double* u, double* v, double* w, const double* uh, const double* p,
                   double* un_out) {
    double* U[3] = {u, v, w};
    for (int c = 0; c < F->dim; ++c) F->uh[c].assign(uh + (size_t)c * F->nv, uh + (size_t)(c + 1) * F->nv);
    std::vector<double> un[3];
    fs_correct(F, U, p, un);
    for (int c = 0; c < F->dim; ++c) std::copy(un[c].begin(), un[c].end(), un_out + (size_t)c * F->nv);
}

void or_fs_last(OrFs* F, int* iters, double* rel, double* abs_) {
    if (iters) *iters = F->last_iters;
    if (rel) *rel = F->last_rel;
    if (abs_) *abs_ = F->last_abs;
}

// Continuity residual Gx u + Gy v (+ Gz w) and its max / mean |.| over
// interior, boundary and all nodes (postprocessing_functions.cpp:192-195,
// 248-251 with calc_max_l1_error, general_functions.cpp:998-1027).
// out = {max_int, mean_int, max_bnd, mean_bnd, max_all, mean_all}
void or_divergence(int nv, int dim, const int* rp, const int* ci, const double* gx,
                   const double* gy, const double* gz, const uint8_t* bnd, const double* u,
                   const double* v, const double* w, double* res, double* out) {
    std::vector<double> a(nv), b(nv), c(nv, 0.0);
    or_spmv(nv, rp, ci, gx, u, a.data());
    or_spmv(nv, rp, ci, gy, v, b.data());
    if (dim == 3) or_spmv(nv, rp, ci, gz, w, c.data());
    double mi = 0, si = 0, mb = 0, sb = 0;
    int ni = 0, nb = 0;
    for (int i = 0; i < nv; ++i) {
        double r = a[i] + b[i];
        if (dim == 3) r = r + c[i];
        if (res) res[i] = r;
        double e = std::fabs(r);
        if (bnd[i]) { nb++; sb += e; mb = std::max(mb, e); }
        else { ni++; si += e; mi = std::max(mi, e); }
    }
    out[0] = mi; out[1] = ni ? si / ni : 0.0;
    out[2] = mb; out[3] = nb ? sb / nb : 0.0;
    out[4] = std::max(mi, mb); out[5] = (si + sb) / nv;
}

}  // extern "C"
