// Host set-up of the smoothed-aggregation hierarchy (declarations and the
// method in amg_build.hpp).
#include "amg_build.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

namespace mp_amg {

namespace {

int thread_count(int requested, long work) {
    int t = requested > 0 ? requested : (int)std::thread::hardware_concurrency();
    if (t < 1) t = 1;
    if (work < 20000) t = 1;
    return std::min(t, 32);
}

// Runs body(r0, r1, tid) over [0, n) split into `t` contiguous ranges.
template <typename F>
void parallel_rows(int n, int t, F&& body) {
    if (t <= 1) {
        body(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    for (int q = 0; q < t; ++q) {
        const int r0 = (int)((long)n * q / t), r1 = (int)((long)n * (q + 1) / t);
        pool.emplace_back([&, r0, r1, q]() { body(r0, r1, q); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

void aggregate(const Csr& A, int k, std::vector<int>& agg, int& n_agg) {
    const int n = A.nrows;
    // directed strength edges: the k largest |a_ij|, j != i (ties: smaller j)
    std::vector<int> dptr(n + 1, 0), dcol;
    dcol.reserve((size_t)n * k);
    std::vector<std::pair<double, int>> cand;
    for (int i = 0; i < n; ++i) {
        cand.clear();
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
            if (A.col[t] != i) cand.emplace_back(-std::fabs(A.val[t]), A.col[t]);
        const int keep = std::min<int>(k, (int)cand.size());
        std::partial_sort(cand.begin(), cand.begin() + keep, cand.end());
        for (int q = 0; q < keep; ++q) dcol.push_back(cand[q].second);
        dptr[i + 1] = (int)dcol.size();
    }
    // symmetrise: adjacency = out-edges and in-edges, sorted, duplicates removed
    std::vector<int> deg(n, 0);
    for (int i = 0; i < n; ++i)
        for (int t = dptr[i]; t < dptr[i + 1]; ++t) {
            deg[i]++;
            deg[dcol[t]]++;
        }
    std::vector<int> sptr(n + 1, 0);
    for (int i = 0; i < n; ++i) sptr[i + 1] = sptr[i] + deg[i];
    std::vector<int> scol(sptr[n]), fill(sptr.begin(), sptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int t = dptr[i]; t < dptr[i + 1]; ++t) {
            const int j = dcol[t];
            scol[fill[i]++] = j;
            scol[fill[j]++] = i;
        }
    std::vector<int> aptr(n + 1, 0), acol;
    acol.reserve(scol.size());
    for (int i = 0; i < n; ++i) {
        auto b = scol.begin() + sptr[i], e = scol.begin() + sptr[i + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        acol.insert(acol.end(), b, e);
        aptr[i + 1] = (int)acol.size();
    }
    agg.assign(n, -1);
    n_agg = 0;
    // pass 1: seeds whose whole neighbourhood is free
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool free_nb = true;
        for (int t = aptr[i]; t < aptr[i + 1] && free_nb; ++t) free_nb = agg[acol[t]] < 0;
        if (!free_nb) continue;
        agg[i] = n_agg;
        for (int t = aptr[i]; t < aptr[i + 1]; ++t) agg[acol[t]] = n_agg;
        n_agg++;
    }
    // pass 2: leftovers join their first aggregated neighbour, else stay alone
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        int a = -1;
        for (int t = aptr[i]; t < aptr[i + 1] && a < 0; ++t)
            if (agg[acol[t]] >= 0) a = agg[acol[t]];
        agg[i] = (a >= 0) ? a : n_agg++;
    }
}

Csr smoothed_prolongator(const Csr& A, const std::vector<int>& agg, int n_agg, double omega,
                         double trunc) {
    const int n = A.nrows;
    Csr P;
    P.nrows = n;
    P.ncols = n_agg;
    P.ptr.assign(n + 1, 0);
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < n; ++i) {
        double diag = 0.0;
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
            if (A.col[t] == i) diag += A.val[t];
        row.clear();
        row.emplace_back(agg[i], 1.0);
        const double f = (diag != 0.0) ? -omega / diag : 0.0;
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) row.emplace_back(agg[A.col[t]], f * A.val[t]);
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                             return a.first < b.first;
                         });
        const size_t first = P.col.size();
        double amax = 0.0;
        for (size_t q = 0; q < row.size();) {
            const int c = row[q].first;
            double s = 0.0;
            while (q < row.size() && row[q].first == c) s += row[q++].second;
            P.col.push_back(c);
            P.val.push_back(s);
            amax = std::max(amax, std::fabs(s));
        }
        if (trunc > 0.0) {
            double all = 0.0, kept = 0.0;
            size_t w = first;
            for (size_t q = first; q < P.col.size(); ++q) {
                all += P.val[q];
                if (std::fabs(P.val[q]) < trunc * amax) continue;
                kept += P.val[q];
                P.col[w] = P.col[q];
                P.val[w++] = P.val[q];
            }
            P.col.resize(w);
            P.val.resize(w);
            if (std::fabs(kept) > 1e-300)
                for (size_t q = first; q < w; ++q) P.val[q] *= all / kept;
        }
        P.ptr[i + 1] = (int)P.col.size();
    }
    return P;
}

Csr sparsify(const Csr& A, double drop) {
    const int n = A.nrows;
    std::vector<double> d(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
            if (A.col[t] == i) d[i] = std::fabs(A.val[t]);
    Csr S;
    S.nrows = n;
    S.ncols = A.ncols;
    S.ptr.assign(n + 1, 0);
    S.col.reserve(A.col.size());
    S.val.reserve(A.val.size());
    for (int i = 0; i < n; ++i) {
        double lumped = 0.0;
        long diag = -1;
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) {
            const int j = A.col[t];
            if (j != i && std::fabs(A.val[t]) < drop * std::sqrt(d[i] * d[j])) {
                lumped += A.val[t];
                continue;
            }
            if (j == i) diag = (long)S.col.size();
            S.col.push_back(j);
            S.val.push_back(A.val[t]);
        }
        if (diag >= 0) S.val[diag] += lumped;
        S.ptr[i + 1] = (int)S.col.size();
    }
    return S;
}

Csr transpose(const Csr& A) {
    Csr T;
    T.nrows = A.ncols;
    T.ncols = A.nrows;
    T.ptr.assign(T.nrows + 1, 0);
    for (int t = 0; t < A.nnz(); ++t) T.ptr[A.col[t] + 1]++;
    for (int r = 0; r < T.nrows; ++r) T.ptr[r + 1] += T.ptr[r];
    T.col.resize(A.nnz());
    T.val.resize(A.nnz());
    std::vector<int> fill(T.ptr.begin(), T.ptr.end() - 1);
    for (int i = 0; i < A.nrows; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) {
            const int d = fill[A.col[t]]++;
            T.col[d] = i;
            T.val[d] = A.val[t];
        }
    return T;
}

// Gustavson row-by-row product; each row's columns come out sorted.  The sum
// for an entry runs over A's row in order, so the result does not depend on
// the thread count.
Csr multiply(const Csr& A, const Csr& B, int threads) {
    if (A.ncols != B.nrows) throw std::runtime_error("amg multiply: shape mismatch");
    const int n = A.nrows;
    const int t = thread_count(threads, A.nnz());
    std::vector<std::vector<int>> cols(t);
    std::vector<std::vector<double>> vals(t);
    std::vector<int> rowlen(n, 0);
    parallel_rows(n, t, [&](int r0, int r1, int q) {
        std::vector<int> mark(B.ncols, -1);
        std::vector<double> acc(B.ncols, 0.0);
        std::vector<int> touched;
        for (int i = r0; i < r1; ++i) {
            touched.clear();
            for (int a = A.ptr[i]; a < A.ptr[i + 1]; ++a) {
                const int k = A.col[a];
                const double av = A.val[a];
                for (int b = B.ptr[k]; b < B.ptr[k + 1]; ++b) {
                    const int c = B.col[b];
                    if (mark[c] != i) {
                        mark[c] = i;
                        acc[c] = 0.0;
                        touched.push_back(c);
                    }
                    acc[c] += av * B.val[b];
                }
            }
            std::sort(touched.begin(), touched.end());
            for (int c : touched) {
                cols[q].push_back(c);
                vals[q].push_back(acc[c]);
            }
            rowlen[i] = (int)touched.size();
        }
    });
    Csr C;
    C.nrows = n;
    C.ncols = B.ncols;
    C.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) C.ptr[i + 1] = C.ptr[i] + rowlen[i];
    C.col.reserve(C.ptr[n]);
    C.val.reserve(C.ptr[n]);
    for (int q = 0; q < t; ++q) {
        C.col.insert(C.col.end(), cols[q].begin(), cols[q].end());
        C.val.insert(C.val.end(), vals[q].begin(), vals[q].end());
    }
    return C;
}

namespace {

// Gauss-Jordan with partial pivoting on [M | I] (m x m, row-major); returns
// M^-1.
std::vector<double> gauss_jordan(std::vector<double> M, int m) {
    std::vector<double> I((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) I[(size_t)i * m + i] = 1.0;
    for (int k = 0; k < m; ++k) {
        int piv = k;
        double best = std::fabs(M[(size_t)k * m + k]);
        for (int i = k + 1; i < m; ++i) {
            const double v = std::fabs(M[(size_t)i * m + k]);
            if (v > best) {
                best = v;
                piv = i;
            }
        }
        if (!(best > 0.0)) throw std::runtime_error("amg: singular coarsest matrix");
        if (piv != k)
            for (int j = 0; j < m; ++j) {
                std::swap(M[(size_t)k * m + j], M[(size_t)piv * m + j]);
                std::swap(I[(size_t)k * m + j], I[(size_t)piv * m + j]);
            }
        const double inv = 1.0 / M[(size_t)k * m + k];
        for (int j = 0; j < m; ++j) {
            M[(size_t)k * m + j] *= inv;
            I[(size_t)k * m + j] *= inv;
        }
        for (int i = 0; i < m; ++i) {
            if (i == k) continue;
            const double f = M[(size_t)i * m + k];
            if (f == 0.0) continue;
            for (int j = 0; j < m; ++j) {
                M[(size_t)i * m + j] -= f * M[(size_t)k * m + j];
                I[(size_t)i * m + j] -= f * I[(size_t)k * m + j];
            }
        }
    }
    return I;
}

}  // namespace

std::vector<double> bordered_inverse(const Csr& A) {
    const int n = A.nrows, m = n + 1;
    // M = [[A, 1], [1^T, 0]]
    std::vector<double> M((size_t)m * m, 0.0);
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) M[(size_t)i * m + A.col[t]] += A.val[t];
    for (int i = 0; i < n; ++i) {
        M[(size_t)i * m + n] = 1.0;
        M[(size_t)n * m + i] = 1.0;
    }
    const std::vector<double> I = gauss_jordan(std::move(M), m);
    std::vector<double> G((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) G[(size_t)i * n + j] = I[(size_t)i * m + j];
    return G;
}

std::vector<double> dense_inverse(const Csr& A) {
    const int n = A.nrows;
    std::vector<double> M((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) M[(size_t)i * n + A.col[t]] += A.val[t];
    return gauss_jordan(std::move(M), n);
}

Hierarchy build(const Csr& A0, const Options& opt, bool bordered) {
    if (A0.nrows != A0.ncols || A0.nrows < 1) throw std::runtime_error("amg: operator not square");
    Hierarchy h;
    h.A.emplace_back();  // level 0 stays with the caller
    const Csr* cur = &A0;
    while (cur->nrows > opt.coarse_max && (int)h.A.size() < opt.max_levels) {
        std::vector<int> agg;
        int na = 0;
        aggregate(*cur, opt.strength_k, agg, na);
        if (na >= cur->nrows) break;  // no coarsening possible
        Csr P = smoothed_prolongator(*cur, agg, na, opt.omega_p, opt.trunc);
        Csr R = transpose(P);
        Csr AP = multiply(*cur, P, opt.threads);
        Csr Ac = multiply(R, AP, opt.threads);
        if (opt.drop > 0.0) Ac = sparsify(Ac, opt.drop);
        h.P.push_back(std::move(P));
        h.R.push_back(std::move(R));
        h.A.push_back(std::move(Ac));
        cur = &h.A.back();
    }
    if (cur->nrows > 4096)
        throw std::runtime_error("amg: coarsest level has " + std::to_string(cur->nrows) +
                                 " rows (dense limit 4096)");
    h.nc = cur->nrows;
    h.coarse_inv = bordered ? bordered_inverse(*cur) : dense_inverse(*cur);
    return h;
}

}  // namespace mp_amg
