// Part (a): batched PHS r^p + polynomial stencil weights, FP64.
//
// One workgroup (256 threads, 4 wavefronts) per stencil.  The saddle-point
// matrix [Phi P; P^T 0] (M = k + m) and the d+1 right-hand sides live in LDS;
// partial-pivot LU runs right-looking with one barrier-separated phase per
// step (pivot search on wave 0, row swap, multipliers, trailing update), then a
// column-oriented back substitution.  The arithmetic order is fixed: each
// matrix entry sees the same fma sequence as the oracle's sequential LU, so
// the weights agree bit for bit with it.
//
// Follows coefficient_computations.cpp:33-328 (shifting_scaling, A and rhs
// assembly, partialPivLu().solve) and cloud.cpp:207-279 (loop over nodes,
// row 0 of the solution is the stencil's weight row).
#include <algorithm>
#include <climits>

#include "mp_internal.hpp"

namespace mp {

namespace {

__device__ __forceinline__ double rpow_d(double r2, int q) {
    if (q >= 1) {
        double v = sqrt(r2);
        for (int s = 0; s < (q - 1) / 2; ++s) v = v * r2;
        return v;
    }
    double v = 1.0 / sqrt(r2);
    for (int s = 0; s < (-q - 1) / 2; ++s) v = v / r2;
    return v;
}

__device__ __forceinline__ double ipow_d(double x, int e) {
    double v = 1.0;
    for (int s = 0; s < e; ++s) v = v * x;
    return v;
}

// Right-hand-side entries of saddle-point row i (columns gx, gy, (gz), lap at
// out[0..dim]) with the oracle's formulas (coefficient_computations.cpp:108-292):
// a stencil row (is_pt) differentiates the PHS kernel between the scaled point
// xi and the scaled centre xc; a polynomial row differentiates the monomial
// with exponents e at xc.  sc holds the per-axis scale factors.
__device__ __forceinline__ void rhs_entries(bool is_pt, int dim, int phs, const double* xi,
                                            const double* xc, const double* sc, const int* e,
                                            double* out) {
    const double p = (double)phs;
    if (is_pt) {
        double r2 = 0.0;
        for (int a = 0; a < dim; ++a) {
            double d = xi[a] - xc[a];
            r2 = r2 + d * d;
        }
        for (int a = 0; a < dim; ++a) {
            double v = (-p) * (xi[a] - xc[a]);
            v = v * rpow_d(r2, phs - 2);
            out[a] = v / sc[a];
        }
        double lp = 0.0;
        if (fabs(r2) > 1e-14) {
            double g[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < dim; ++a) {
                double d = xi[a] - xc[a];
                double d2 = d * d;
                double q4 = rpow_d(r2, phs - 4);
                double v = (p * p) * d2 * q4;
                v = v - (2.0 * p) * d2 * q4;
                v = v + p * rpow_d(r2, phs - 2);
                g[a] = v / (sc[a] * sc[a]);
            }
            lp = g[0] + g[1];
            if (dim == 3) lp = lp + g[2];
        }
        out[dim] = lp;
        return;
    }
    const double x0 = xc[0], x1 = xc[1], x2 = (dim == 3) ? xc[2] : 0.0;
    for (int a = 0; a < dim; ++a) {
        double v = 0.0;
        if (e[a] > 0) {
            if (a == 1) {
                v = ipow_d(x0, e[0]);
                v = v * (double)e[1];
                v = v * ipow_d(x1, e[1] - 1);
                if (dim == 3) v = v * ipow_d(x2, e[2]);
            } else if (a == 0) {
                v = (double)e[0] * ipow_d(x0, e[0] - 1);
                v = v * ipow_d(x1, e[1]);
                if (dim == 3) v = v * ipow_d(x2, e[2]);
            } else {
                v = (double)e[2] * ipow_d(x2, e[2] - 1);
                v = v * ipow_d(x0, e[0]);
                v = v * ipow_d(x1, e[1]);
            }
            v = v / sc[a];
        }
        out[a] = v;
    }
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (e[0] > 1) {
        double v = (double)e[0] * ((double)e[0] - 1.0) * ipow_d(x0, e[0] - 2);
        v = v * ipow_d(x1, e[1]);
        if (dim == 3) v = v * ipow_d(x2, e[2]);
        g0 = v / (sc[0] * sc[0]);
    }
    if (e[1] > 1) {
        double v = (double)e[1] * ((double)e[1] - 1.0) * ipow_d(x1, e[1] - 2);
        v = v * ipow_d(x0, e[0]);
        if (dim == 3) v = v * ipow_d(x2, e[2]);
        g1 = v / (sc[1] * sc[1]);
    }
    if (dim == 3 && e[2] > 1) {
        double v = (double)e[2] * ((double)e[2] - 1.0) * ipow_d(x2, e[2] - 2);
        v = v * ipow_d(x0, e[0]);
        v = v * ipow_d(x1, e[1]);
        g2 = v / (sc[2] * sc[2]);
    }
    double lp = g0 + g1;
    if (dim == 3) lp = lp + g2;
    out[dim] = lp;
}

__global__ __launch_bounds__(256) void weights_fp64_kernel(
    int dim, int phs, int m, const int* __restrict__ exps, SellView sv,
    const double* __restrict__ xyz, int npad, double* __restrict__ gx, double* __restrict__ gy,
    double* __restrict__ gz, double* __restrict__ lap, double* __restrict__ cond,
    const int* __restrict__ rows, const int* __restrict__ only, int t0, double* gwork,
    long gstride) {
    extern __shared__ double lds[];
    __shared__ double s_mn[3], s_sc[3];
    __shared__ double s_anorm, s_red[4];
    __shared__ int s_piv, s_jm;
    __shared__ int s_e[MAX_POLY * 3];

    // stencil t of the list (or row t); the work arrays in LDS, or for
    // systems too large for it in this block's share of a global work area
    // (the same operations in the same order either way)
    const int t = t0 + (int)blockIdx.x;
    if (only && !only[t]) return;  // a stencil the register kernel finished
    const int r = rows ? rows[t] : t;
    if (r >= sv.nv || sv.len[r] == 0) return;  // padding rows carry no stencil
    const int tid = threadIdx.x;
    const int k = sv.len[r];
    const int M = k + m;
    const int LDA = M + 1;
    const int nr = dim + 1;
    double* A = gwork ? gwork + (size_t)blockIdx.x * (size_t)gstride : lds;
    double* R = A + M * LDA;
    double* xs = R + M * 4;
    const int base = sell_base(sv.sptr, r);

    for (int i = tid; i < m * dim; i += blockDim.x) s_e[i] = exps[i];
    for (int t = tid; t < k; t += blockDim.x) {
        const int c = sv.col[base + sell_off(t)];
        for (int a = 0; a < dim; ++a) xs[t * 3 + a] = xyz[(size_t)a * npad + c];
    }
    __syncthreads();
    if (tid < dim) {
        double mn = xs[tid], mx = xs[tid];
        for (int t = 0; t < k; ++t) {
            double v = xs[t * 3 + tid];
            if (mn > v) mn = v;
            if (mx < v) mx = v;
        }
        s_mn[tid] = mn;
        s_sc[tid] = mx - mn;
    }
    __syncthreads();
    for (int t = tid; t < k * dim; t += blockDim.x) {
        const int j = t / dim, a = t - j * dim;
        xs[j * 3 + a] = (xs[j * 3 + a] - s_mn[a]) / s_sc[a];
    }
    __syncthreads();

    // ---- assemble A (M x M) ----
    for (int idx = tid; idx < M * M; idx += blockDim.x) {
        const int i = idx / M, j = idx - i * M;
        double v = 0.0;
        if (i < k && j < k) {
            if (i != j) {
                double s = 0.0;
                for (int a = 0; a < dim; ++a) {
                    double d = xs[i * 3 + a] - xs[j * 3 + a];
                    s = s + d * d;
                }
                v = rpow_d(s, phs);
            }
        } else if (i < k || j < k) {
            const int pt = (i < k) ? i : j;
            const int t = (i < k) ? j - k : i - k;
            v = ipow_d(xs[pt * 3 + 0], s_e[t * dim + 0]);
            for (int a = 1; a < dim; ++a) v = v * ipow_d(xs[pt * 3 + a], s_e[t * dim + a]);
        }
        A[i * LDA + j] = v;
    }
    // ---- right-hand sides: gx, gy, (gz), lap ----
    for (int i = tid; i < M; i += blockDim.x)
        rhs_entries(i < k, dim, phs, xs + 3 * (i < k ? i : 0), xs, s_sc,
                    s_e + (i < k ? 0 : i - k) * dim, R + 4 * i);
    __syncthreads();

    // ---- 1-norm of A (for the condition estimate) ----
    int* perm = (int*)xs;  // the scaled coordinates are no longer needed
    if (cond) {
        if (tid == 0) s_anorm = 0.0;
        __syncthreads();
        for (int j = tid; j < M; j += blockDim.x) {
            double sum = 0.0;
            for (int i = 0; i < M; ++i) sum += fabs(A[i * LDA + j]);
            atomicMax((unsigned long long*)&s_anorm, (unsigned long long)__double_as_longlong(sum));
        }
    }
    __syncthreads();
    for (int i = tid; i < M; i += blockDim.x) perm[i] = i;
    __syncthreads();

    // ---- LU with partial pivoting, right-looking ----
    for (int c = 0; c < M; ++c) {
        if (tid < 64) {
            double best = -1.0;
            int bi = M;
            for (int i = c + tid; i < M; i += 64) {
                double v = fabs(A[i * LDA + c]);
                if (v > best) {
                    best = v;
                    bi = i;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                double ob = __shfl_xor(best, off);
                int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (tid == 0) s_piv = bi;
        }
        __syncthreads();
        const int piv = s_piv;
        if (piv != c) {
            if (tid == 0) {
                const int tp = perm[c];
                perm[c] = perm[piv];
                perm[piv] = tp;
            }
            for (int j = tid; j < M + nr; j += blockDim.x) {
                if (j < M) {
                    double t0 = A[c * LDA + j];
                    A[c * LDA + j] = A[piv * LDA + j];
                    A[piv * LDA + j] = t0;
                } else {
                    const int q = j - M;
                    double t0 = R[c * 4 + q];
                    R[c * 4 + q] = R[piv * 4 + q];
                    R[piv * 4 + q] = t0;
                }
            }
            __syncthreads();
        }
        const double d = A[c * LDA + c];
        if (d != 0.0) {
            for (int i = c + 1 + tid; i < M; i += blockDim.x) A[i * LDA + c] = A[i * LDA + c] / d;
            __syncthreads();
            const int nrow = M - c - 1;
            const int ncol = M - c - 1 + nr;
            for (int idx = tid; idx < nrow * ncol; idx += blockDim.x) {
                const int ii = idx / ncol;
                const int jj = idx - ii * ncol;
                const int i = c + 1 + ii;
                const double l = A[i * LDA + c];
                if (jj < nrow) {
                    const int j = c + 1 + jj;
                    A[i * LDA + j] = fma(-l, A[c * LDA + j], A[i * LDA + j]);
                } else {
                    const int q = jj - nrow;
                    R[i * 4 + q] = fma(-l, R[c * 4 + q], R[i * 4 + q]);
                }
            }
        }
        __syncthreads();
    }
    // ---- back substitution (column oriented) ----
    for (int c = M - 1; c >= 0; --c) {
        if (tid < nr) R[c * 4 + tid] = R[c * 4 + tid] / A[c * LDA + c];
        __syncthreads();
        for (int idx = tid; idx < c * nr; idx += blockDim.x) {
            const int i = idx / nr, q = idx - i * nr;
            R[i * 4 + q] = fma(-A[i * LDA + c], R[c * 4 + q], R[i * 4 + q]);
        }
        __syncthreads();
    }
    for (int t = tid; t < k; t += blockDim.x) {
        const size_t slot = (size_t)base + (size_t)sell_off(t);
        gx[slot] = R[t * 4 + 0];
        gy[slot] = R[t * 4 + 1];
        if (dim == 3) gz[slot] = R[t * 4 + 2];
        lap[slot] = R[t * 4 + dim];
    }
    if (!cond) return;
    __syncthreads();
    // ---- Hager (1984) / Higham (1988) estimate of ||A^-1||_1 from the LU factors ----
    double* w = R;          // work vectors reuse the right-hand-side storage
    double* xv = R + M;
    double* yv = R + 2 * M;
    bool singular = false;
    for (int c = 0; c < M; ++c) singular |= (A[c * LDA + c] == 0.0);
    if (singular) {
        if (tid == 0) cond[r] = __builtin_inf();
        return;
    }
    // w = A^-1 v  (v in xv): w = P v, forward with unit L, backward with U
    auto solve = [&](const double* v, double* out) {
        for (int i = tid; i < M; i += blockDim.x) out[i] = v[perm[i]];
        __syncthreads();
        for (int c = 0; c < M; ++c) {
            for (int i = c + 1 + tid; i < M; i += blockDim.x) out[i] = fma(-A[i * LDA + c], out[c], out[i]);
            __syncthreads();
        }
        for (int c = M - 1; c >= 0; --c) {
            if (tid == 0) out[c] = out[c] / A[c * LDA + c];
            __syncthreads();
            for (int i = tid; i < c; i += blockDim.x) out[i] = fma(-A[i * LDA + c], out[c], out[i]);
            __syncthreads();
        }
    };
    // out = A^-T v: U^T then L^T (column oriented), then un-permute
    auto solve_t = [&](const double* v, double* tmpv, double* out) {
        for (int i = tid; i < M; i += blockDim.x) tmpv[i] = v[i];
        __syncthreads();
        for (int c = 0; c < M; ++c) {
            if (tid == 0) tmpv[c] = tmpv[c] / A[c * LDA + c];
            __syncthreads();
            for (int i = c + 1 + tid; i < M; i += blockDim.x) tmpv[i] = fma(-A[c * LDA + i], tmpv[c], tmpv[i]);
            __syncthreads();
        }
        for (int c = M - 1; c >= 0; --c) {
            for (int i = tid; i < c; i += blockDim.x) tmpv[i] = fma(-A[c * LDA + i], tmpv[c], tmpv[i]);
            __syncthreads();
        }
        for (int i = tid; i < M; i += blockDim.x) out[perm[i]] = tmpv[i];
        __syncthreads();
    };
    auto l1 = [&](const double* v) {  // sequential sum on thread 0, broadcast
        if (tid == 0) {
            double sum = 0.0;
            for (int i = 0; i < M; ++i) sum += fabs(v[i]);
            s_red[0] = sum;
        }
        __syncthreads();
        double res = s_red[0];
        __syncthreads();
        return res;
    };
    for (int i = tid; i < M; i += blockDim.x) xv[i] = 1.0 / (double)M;
    __syncthreads();
    double est = 0.0;
    int jlast = -1;
    for (int itn = 0; itn < 5; ++itn) {
        solve(xv, yv);
        est = fmax(est, l1(yv));
        for (int i = tid; i < M; i += blockDim.x) w[i] = (yv[i] >= 0.0) ? 1.0 : -1.0;
        __syncthreads();
        solve_t(w, yv, w);
        if (tid == 0) {
            int jm = 0;
            double zx = 0.0;
            for (int i = 0; i < M; ++i) {
                if (fabs(w[i]) > fabs(w[jm])) jm = i;
                zx += w[i] * xv[i];
            }
            s_jm = (fabs(w[jm]) <= zx || jm == jlast) ? -1 : jm;
        }
        __syncthreads();
        const int jm = s_jm;
        if (jm < 0) break;
        jlast = jm;
        for (int i = tid; i < M; i += blockDim.x) xv[i] = (i == jm) ? 1.0 : 0.0;
        __syncthreads();
    }
    for (int i = tid; i < M; i += blockDim.x)
        xv[i] = ((i % 2) ? -1.0 : 1.0) * (1.0 + (double)i / (double)(M > 1 ? M - 1 : 1));
    __syncthreads();
    solve(xv, yv);
    est = fmax(est, 2.0 * l1(yv) / (3.0 * (double)M));
    if (tid == 0) cond[r] = s_anorm * est;
}


// ---------------------------------------------------------------------------
// FP64 weights with the saddle-point matrix in registers: one wavefront per
// stencil (two for 64 < M <= 96), thread i holding row i.  At pivot step c a
// row keeps its columns c..M-1 in a[0..M-1-c]: the trailing update writes a[j]
// from a[j+1], so the row shifts by one column per step and every register
// index is a compile-time constant.  The update width is rounded up to a
// multiple of 8 (the entries past the row's end are never read back).  The
// pivot row is published through LDS, where it stays as row c of U (packed,
// row c at c*M - c(c-1)/2) for the back substitution, and its forward-
// eliminated right-hand sides stay in Y[c].  Every entry sees the oracle's
// roundings in the oracle's order: the same division for each multiplier,
// fma(-l, u, a) for each update, the pivot chosen as the first maximum in
// position order (a NaN on the diagonal position keeps it), and the column-
// oriented back substitution.  The weights are therefore bit-identical to the
// oracle's and to the LDS kernel's.

__device__ __forceinline__ double readlane_d(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// d = a * b + c, one rounding, as the three-address v_fma_f64: the result may
// go to any register.  (The compiler's own fma becomes the two-address
// v_fmac_f64, which writes over c; for the shifted rows below that register is
// the neighbour's, and every step would end by copying the row back.)
__device__ __forceinline__ double fma3(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// Maximum of a double over the wavefront (no NaN among the inputs), through
// DPP row shifts and broadcasts: lane 63 ends with the maximum, which every
// lane receives.  Max is exact and order-free, so any combination order gives
// the same value.
__device__ __forceinline__ double wave_max_d(double v) {
    int hi = __double2hiint(v), lo = __double2loint(v);
#define MP_DPP_MAX(ctrl, rows)                                                    \
    {                                                                             \
        const int th = __builtin_amdgcn_update_dpp(hi, hi, ctrl, rows, 0xf, false); \
        const int tl = __builtin_amdgcn_update_dpp(lo, lo, ctrl, rows, 0xf, false); \
        v = fmax(v, __hiloint2double(th, tl));                                    \
        hi = __double2hiint(v);                                                   \
        lo = __double2loint(v);                                                   \
    }
    MP_DPP_MAX(0x111, 0xf)  // row_shr:1
    MP_DPP_MAX(0x112, 0xf)  // row_shr:2
    MP_DPP_MAX(0x114, 0xf)  // row_shr:4
    MP_DPP_MAX(0x118, 0xf)  // row_shr:8: lane 15 of each row holds the row's maximum
    MP_DPP_MAX(0x142, 0xa)  // row_bcast:15 into rows 1 and 3
    MP_DPP_MAX(0x143, 0xc)  // row_bcast:31 into rows 2 and 3: lane 63 holds the maximum
#undef MP_DPP_MAX
    return __hiloint2double(__builtin_amdgcn_readlane(hi, 63), __builtin_amdgcn_readlane(lo, 63));
}

// Minimum and maximum over the lanes that hold a point, each equal to the
// oracle's left-to-right scan (a later value replaces the current one only when
// strictly smaller / larger): the extreme value by DPP reductions, then the
// exact bits of the first lane (lowest index) holding an equal value, which is
// the one the scan keeps among equal values such as -0 and +0.
__device__ __forceinline__ void ordered_minmax(double& mn, double& mx, int& has, int lane) {
    (void)lane;
    const double big = __builtin_inf();
    const double x = mn;
    const double vmax = wave_max_d(has ? x : -big);
    const double vmin = -wave_max_d(has ? -x : -big);
    const unsigned long long bmax = __ballot(has && x == vmax);
    const unsigned long long bmin = __ballot(has && x == vmin);
    mx = bmax ? readlane_d(x, __ffsll((long long)bmax) - 1) : vmax;
    mn = bmin ? readlane_d(x, __ffsll((long long)bmin) - 1) : vmin;
    has = __ballot(has) != 0ull;
}

// LDS doubles the register-resident kernel needs for a stencil of M = k + m.
__host__ __device__ inline int wave_lds_doubles(int M, int k, int m) {
    const int lu = M * (M + 1) / 2 + 8 + 4 * M;  // packed U (+ overflow) and Y
    const int asm_ = 4 * k + m * k;             // scaled coordinates and monomials
    return lu > asm_ ? lu : asm_;
}

// The pivot steps of one wavefront's rows.  phase<W> runs the steps whose
// trailing width M-1-c lies in (W-8, W] with a fixed update width W, then hands
// over to phase<W-8>: each loop body has one static width, so the shifted row
// stays in the same registers from step to step.
template <int NW, int AW>
struct WaveLu {
    double (&a)[AW + 1];
    double (&r)[4];
    double* U;
    double* Y;
    const double* zero;  // AW + 1 zeros
    double* s_cv;
    int* s_cp;
    int* s_cn;
    int M, pos, lane, wv;
    bool valid;
    int ub = 0;         // start of U row c
    bool zero_pivot = false;  // a single wavefront met an exactly zero pivot (stencil redone)

    // One wavefront (M <= 64): no LDS inside the factorisation.  The pivot is
    // found with a DPP maximum and a ballot, its row is broadcast with
    // readlane, and only the rows below the pivot are updated (exec mask), so
    // a pivot row keeps its U row in registers until the end (store_u).
    template <int W>
    __device__ __forceinline__ void step1(int c) {
        const double av = fabs(a[0]);
        const bool act = valid && pos >= c;
        const bool isn = !(av == av);
        const double cv = (act && !isn) ? av : -1.0;
        const double mxv = wave_max_d(cv);
        const unsigned long long tie = __ballot(act && cv == mxv && mxv >= 0.0);
        int piv;
        if (__ballot(valid && pos == c && isn) != 0ull || tie == 0ull) {
            piv = c;  // a NaN on the diagonal position keeps it (nothing compares greater)
        } else if (__popcll(tie) == 1) {
            piv = __builtin_amdgcn_readlane(pos, __ffsll((long long)tie) - 1);
        } else {  // exact ties: the first in position order
            int cp = (act && cv == mxv) ? pos : INT_MAX;
            for (int off = 32; off > 0; off >>= 1) cp = min(cp, __shfl_xor(cp, off));
            piv = __builtin_amdgcn_readfirstlane(cp);
        }
        if (pos == c) pos = piv;
        else if (pos == piv) pos = c;
        const int pl = __ffsll((long long)__ballot(pos == c)) - 1;  // the pivot row's lane
        // The pivot row reaches the other rows two ways at once: its first
        // L+1 entries through LDS (one lane writes, every lane reads the same
        // address), the rest by readlane, which balances the LDS and vector
        // pipes (LDS alone or readlane alone each take about 1.5x longer).
        constexpr int L = (3 * W / 4) & ~1;
        if (pos == c) {
#pragma unroll
            for (int j = 0; j <= L; ++j) U[ub + j] = a[j];
        }
        __syncthreads();
        const double* u = U + ub;
        const double d = readlane_d(a[0], pl);
        if (d == 0.0) {  // a singular column (the oracle skips it): left to the LDS kernel
            zero_pivot = true;
            return;
        }
        if (pos > c) {
            const double nl = -(a[0] / d);
#pragma unroll
            for (int j = 0; j < L; ++j) a[j] = fma3(nl, u[j + 1], a[j + 1]);
#pragma unroll
            for (int j = L; j < W; ++j) a[j] = fma3(nl, readlane_d(a[j + 1], pl), a[j + 1]);
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = fma3(nl, readlane_d(r[q], pl), r[q]);
        }
        ub += M - c;
    }

    // After the last step every row holds its U row (columns pos..M-1 in
    // a[0..M-1-pos]) and its forward-eliminated right-hand sides.
    __device__ __forceinline__ void store_u() {
        if (!valid) return;
        const int b = pos * M - pos * (pos - 1) / 2;
#pragma unroll
        for (int j = 0; j < AW; ++j)
            if (j < M - pos) U[b + j] = a[j];
#pragma unroll
        for (int q = 0; q < 4; ++q) Y[4 * pos + q] = r[q];
    }

    template <int W>
    __device__ __forceinline__ void step(int c) {
        if constexpr (NW == 1) {
            step1<W>(c);
            return;
        }
        // pivot: the first maximum of |column c| in position order
        const double av = fabs(a[0]);
        const bool act = valid && pos >= c;
        const bool isn = !(av == av);
        const double cv = (act && !isn) ? av : -1.0;
        double mxv = cv;
        for (int off = 32; off > 0; off >>= 1) mxv = fmax(mxv, __shfl_xor(mxv, off));
        int cp = (cv == mxv && mxv >= 0.0) ? pos : INT_MAX;
        const unsigned long long tie = __ballot(cp != INT_MAX);
        if (__popcll(tie) == 1) {
            cp = __builtin_amdgcn_readlane(cp, __ffsll((long long)tie) - 1);
        } else {
            for (int off = 32; off > 0; off >>= 1) cp = min(cp, __shfl_xor(cp, off));
        }
        int nanc = __any(valid && pos == c && isn) ? 1 : 0;
        if (NW == 2) {
            if (lane == 0) {
                s_cv[wv] = mxv;
                s_cp[wv] = cp;
                s_cn[wv] = nanc;
            }
            __syncthreads();
            const double v0 = s_cv[0], v1 = s_cv[1];
            const int p0 = s_cp[0], p1 = s_cp[1];
            nanc = s_cn[0] | s_cn[1];
            mxv = v0 > v1 ? v0 : v1;
            cp = (v0 == v1) ? min(p0, p1) : (v0 > v1 ? p0 : p1);
        }
        const int piv = (nanc || mxv < 0.0) ? c : cp;
        if (pos == c) pos = piv;
        else if (pos == piv) pos = c;
        // the pivot row becomes row c of U (entries past its end are overwritten later)
        if (pos == c) {
#pragma unroll
            for (int j = 0; j <= W; ++j) U[ub + j] = a[j];
#pragma unroll
            for (int q = 0; q < 4; ++q) Y[4 * c + q] = r[q];
        }
        __syncthreads();
        // A zero pivot leaves the rows below untouched (the oracle skips the
        // column): multiplier -0 against a row of +0 gives x + (-0) = x exactly
        // for every x, so the same instructions run in both cases.
        const double d = U[ub];
        const bool sing = (d == 0.0);
        const double nl = sing ? -0.0 : -(a[0] / d);
        const double* u = sing ? zero : U + ub;
        const double* y = sing ? zero : Y + 4 * c;
#pragma unroll
        for (int j = 0; j < W; ++j) a[j] = fma3(nl, u[j + 1], a[j + 1]);
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = fma3(nl, y[q], r[q]);
        ub += M - c;
    }

    template <int W>
    __device__ __forceinline__ void phase(int& c) {
        for (; c < M && M - 1 - c > W - 8 && !zero_pivot; ++c) step<W>(c);
        if constexpr (W > 0) phase<W - 8>(c);
    }
};

template <int NW, int AW>
__global__ __launch_bounds__(64 * NW) void weights_wave_kernel(
    int dim, int phs, int m, const int* __restrict__ exps, SellView sv,
    const double* __restrict__ xyz, int npad, double* __restrict__ gx, double* __restrict__ gy,
    double* __restrict__ gz, double* __restrict__ lap, const int* __restrict__ rows,
    int* __restrict__ redo) {
    extern __shared__ double lds[];
    __shared__ double s_mm[NW][6];
    __shared__ int s_has[NW];
    __shared__ double s_sc[3], s_mn[3];
    __shared__ double s_cv[NW];
    __shared__ int s_cp[NW], s_cn[NW];
    __shared__ int s_e[WAVE_MAX_POLY * 3];
    __shared__ double s_zero[AW + 1];

    const int r = rows[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int k = sv.len[r];
    const int M = k + m;
    const int base = sell_base(sv.sptr, r);
    double* U = lds;                          // packed rows of U
    double* Y = lds + M * (M + 1) / 2 + 8;    // forward-eliminated right-hand sides [M][4]
    double* xs = lds;                         // assembly: scaled coordinates [k][4]
    double* P = lds + 4 * k;                  // assembly: monomial t at point j, P[t*k + j]

    for (int i = tid; i < m * dim; i += 64 * NW) s_e[i] = exps[i];
    // ---- shifting_scaling: thread t holds point t ----
    double xp[3] = {0.0, 0.0, 0.0};
    int has = tid < k;
    if (has) {
        const int cidx = sv.col[base + sell_off(tid)];
        for (int a = 0; a < dim; ++a) xp[a] = xyz[(size_t)a * npad + cidx];
    }
    for (int a = 0; a < dim; ++a) {
        double mn = xp[a], mx = xp[a];
        int h = has;
        ordered_minmax(mn, mx, h, lane);
        if (lane == 0) {
            s_mm[wv][2 * a] = mn;
            s_mm[wv][2 * a + 1] = mx;
            s_has[wv] = h;
        }
    }
    __syncthreads();
    if (tid < dim) {
        double mn = s_mm[0][2 * tid], mx = s_mm[0][2 * tid + 1];
        if constexpr (NW == 2) {
            if (s_has[1]) {
                const double mn1 = s_mm[1][2 * tid], mx1 = s_mm[1][2 * tid + 1];
                if (mn > mn1) mn = mn1;
                if (mx < mx1) mx = mx1;
            }
        }
        s_mn[tid] = mn;
        s_sc[tid] = mx - mn;
    }
    __syncthreads();
    if (has) {
        for (int a = 0; a < dim; ++a) {
            xp[a] = (xp[a] - s_mn[a]) / s_sc[a];
            xs[tid * 4 + a] = xp[a];
        }
        for (int t = 0; t < m; ++t) {
            const int* e = s_e + t * dim;
            double v = ipow_d(xp[0], e[0]);
            for (int a = 1; a < dim; ++a) v = v * ipow_d(xp[a], e[a]);
            P[t * k + tid] = v;
        }
    }
    __syncthreads();

    // ---- row i of A = [Phi P; P^T 0] and its right-hand sides ----
    const int i = tid;
    const bool valid = i < M;
    const int ip = i < k ? i : 0;                     // the row's point (stencil rows)
    const int it = (i >= k && i < M) ? i - k : 0;     // the row's monomial (polynomial rows)
    const double xi0 = xs[ip * 4], xi1 = xs[ip * 4 + 1], xi2 = xs[ip * 4 + 2];
    double a[AW + 1];
#pragma unroll
    for (int j = 0; j < AW; ++j) {
        double v = 0.0;
        if (j < M) {
            if (j < k) {
                const double d0 = xi0 - xs[j * 4], d1 = xi1 - xs[j * 4 + 1];
                double s = 0.0;
                s = s + d0 * d0;
                s = s + d1 * d1;
                if (dim == 3) {
                    const double d2 = xi2 - xs[j * 4 + 2];
                    s = s + d2 * d2;
                }
                const double rb = (i == j) ? 0.0 : rpow_d(s, phs);
                v = (i < k) ? rb : P[it * k + j];
            } else {
                v = (i < k) ? P[(j - k) * k + ip] : 0.0;
            }
        }
        a[j] = v;
    }
    a[AW] = 0.0;
    double rr[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid) rhs_entries(i < k, dim, phs, xs + 4 * ip, xs, s_sc, s_e + it * dim, rr);
    __syncthreads();  // the assembly tables are overwritten by U from here on

    // ---- LU with partial pivoting, right-looking ----
    for (int j = tid; j <= AW; j += 64 * NW) s_zero[j] = 0.0;
    __syncthreads();
    WaveLu<NW, AW> lu{a, rr, U, Y, s_zero, s_cv, s_cp, s_cn, M, valid ? i : (1 << 20), lane, wv, valid};
    int c = 0;
    lu.template phase<AW>(c);
    if (lu.zero_pivot) {  // uniform: the whole wavefront saw the zero pivot
        if (tid == 0) redo[blockIdx.x] = 1;
        return;
    }
    if (NW == 1) lu.store_u();
    __syncthreads();

    // ---- back substitution, column oriented (wavefront 0, positions lane and lane+64) ----
    if (wv != 0) return;
    const int t0 = lane, t1 = lane + 64;
    double y0[4] = {0.0, 0.0, 0.0, 0.0}, y1[4] = {0.0, 0.0, 0.0, 0.0};
    if (t0 < M)
        for (int q = 0; q < 4; ++q) y0[q] = Y[4 * t0 + q];
    if (NW == 2 && t1 < M)
        for (int q = 0; q < 4; ++q) y1[q] = Y[4 * t1 + q];
    const int b0 = t0 * M - t0 * (t0 - 1) / 2 - t0;  // U[t][c] = U[b_t + c]
    const int b1 = t1 * M - t1 * (t1 - 1) / 2 - t1;
    for (int c = M - 1; c >= 0; --c) {
        const double d = U[c * M - c * (c - 1) / 2];
        double x[4];
        if (NW == 1 || c < 64) {
#pragma unroll
            for (int q = 0; q < 4; ++q) x[q] = readlane_d(y0[q], c & 63) / d;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) x[q] = readlane_d(y1[q], c & 63) / d;
        }
        if (t0 < c) {
            const double u = U[b0 + c];
#pragma unroll
            for (int q = 0; q < 4; ++q) y0[q] = fma(-u, x[q], y0[q]);
        } else if (t0 == c) {
#pragma unroll
            for (int q = 0; q < 4; ++q) y0[q] = x[q];
        }
        if (NW == 2) {
            if (t1 < c) {
                const double u = U[b1 + c];
#pragma unroll
                for (int q = 0; q < 4; ++q) y1[q] = fma(-u, x[q], y1[q]);
            } else if (t1 == c) {
#pragma unroll
                for (int q = 0; q < 4; ++q) y1[q] = x[q];
            }
        }
    }
    auto put = [&](int t, const double (&y)[4]) {
        const size_t slot = (size_t)base + (size_t)sell_off(t);
        gx[slot] = y[0];
        gy[slot] = y[1];
        if (dim == 3) gz[slot] = y[2];
        lap[slot] = y[dim];
    };
    if (t0 < k) put(t0, y0);
    if (NW == 2 && t1 < k) put(t1, y1);
}

// Entry (i, j) of the saddle-point matrix from the scaled coordinates, with the
// arithmetic of the FP64 assembly above.
__device__ __forceinline__ double saddle_entry(int i, int j, int k, int dim, int phs,
                                               const double* xs, const int* e) {
    if (i < k && j < k) {
        if (i == j) return 0.0;
        double s = 0.0;
        for (int a = 0; a < dim; ++a) {
            double d = xs[i * 3 + a] - xs[j * 3 + a];
            s = s + d * d;
        }
        return rpow_d(s, phs);
    }
    if (i >= k && j >= k) return 0.0;
    const int pt = (i < k) ? i : j;
    const int t = (i < k) ? j - k : i - k;
    double v = ipow_d(xs[pt * 3 + 0], e[t * dim + 0]);
    for (int a = 1; a < dim; ++a) v = v * ipow_d(xs[pt * 3 + a], e[t * dim + a]);
    return v;
}

// FP32 LU + FP64 iterative refinement.  The LU factors (FP32) and the FP64
// right-hand sides live in LDS; each refinement forms r = b - A x in FP64 with
// A's entries recomputed from the coordinates, solves the correction with the
// FP32 factors and adds it in FP64.  Stops once the correction reaches the
// FP64 floor; a stencil that has not converged after `max_ref` rounds is
// flagged for the FP64 kernel.
constexpr int MIX_THREADS = 256;  // four wavefronts per stencil (one: 8 per CU by LDS, 43 % slower)
__global__ __launch_bounds__(MIX_THREADS) void weights_mixed_kernel(
    int dim, int phs, int m, const int* __restrict__ exps, SellView sv,
    const double* __restrict__ xyz, int npad, double* __restrict__ gx, double* __restrict__ gy,
    double* __restrict__ gz, double* __restrict__ lap, const int* __restrict__ rows,
    int* __restrict__ flag, int max_ref) {
    extern __shared__ double lds[];
    __shared__ double s_mn[3], s_sc[3];
    __shared__ float s_bestf;
    __shared__ int s_piv, s_conv;
    __shared__ double s_red[2];
    __shared__ int s_e[WAVE_MAX_POLY * 3];
    const int r = rows[blockIdx.x];
    const int tid = threadIdx.x;
    const int k = sv.len[r];
    const int M = k + m;
    const int LDA = M + 1;
    const int nr = dim + 1;
    double* B = lds;                       // FP64 right-hand sides [M][4]
    double* X = B + M * 4;                 // FP64 solution [M][4]
    double* xs = X + M * 4;                // scaled coordinates [k][3]
    float* Af = (float*)(xs + k * 3);      // FP32 LU factors [M][LDA]
    float* W = Af + M * LDA;               // FP32 work [M][4]
    int* perm = (int*)(W + M * 4);         // row permutation [M]
    const int base = sell_base(sv.sptr, r);
    (void)s_bestf;

    for (int i = tid; i < m * dim; i += blockDim.x) s_e[i] = exps[i];
    for (int t = tid; t < k; t += blockDim.x) {
        const int c = sv.col[base + sell_off(t)];
        for (int a = 0; a < dim; ++a) xs[t * 3 + a] = xyz[(size_t)a * npad + c];
    }
    __syncthreads();
    if (tid < dim) {
        double mn = xs[tid], mx = xs[tid];
        for (int t = 0; t < k; ++t) {
            double v = xs[t * 3 + tid];
            if (mn > v) mn = v;
            if (mx < v) mx = v;
        }
        s_mn[tid] = mn;
        s_sc[tid] = mx - mn;
    }
    __syncthreads();
    for (int t = tid; t < k * dim; t += blockDim.x) {
        const int j = t / dim, a = t - j * dim;
        xs[j * 3 + a] = (xs[j * 3 + a] - s_mn[a]) / s_sc[a];
    }
    __syncthreads();
    for (int i = tid >> 6; i < M; i += MIX_THREADS / 64)
        for (int j = tid & 63; j < M; j += 64)
            Af[i * LDA + j] = (float)saddle_entry(i, j, k, dim, phs, xs, s_e);
    // FP64 right-hand sides (same formulas as the FP64 kernel)
    for (int i = tid; i < M; i += blockDim.x) {
        for (int q = 0; q < 4; ++q) B[i * 4 + q] = 0.0;
        rhs_entries(i < k, dim, phs, xs + 3 * (i < k ? i : 0), xs, s_sc,
                    s_e + (i < k ? 0 : i - k) * dim, B + 4 * i);
        perm[i] = i;
    }
    __syncthreads();
    // FP32 LU with partial pivoting
    bool singular = false;
    for (int c = 0; c < M; ++c) {
        if (tid < 64) {
            float best = -1.0f;
            int bi = M;
            for (int i = c + tid; i < M; i += 64) {
                float v = fabsf(Af[i * LDA + c]);
                if (v > best) {
                    best = v;
                    bi = i;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                float ob = __shfl_xor(best, off);
                int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (tid == 0) s_piv = bi;
        }
        __syncthreads();
        const int piv = s_piv;
        if (piv != c) {
            if (tid == 0) {
                const int tp = perm[c];
                perm[c] = perm[piv];
                perm[piv] = tp;
            }
            for (int j = tid; j < M; j += blockDim.x) {
                float t0 = Af[c * LDA + j];
                Af[c * LDA + j] = Af[piv * LDA + j];
                Af[piv * LDA + j] = t0;
            }
            __syncthreads();
        }
        const float d = Af[c * LDA + c];
        if (d == 0.0f) {
            singular = true;
            __syncthreads();
            continue;
        }
        for (int i = c + 1 + tid; i < M; i += blockDim.x) Af[i * LDA + c] = Af[i * LDA + c] / d;
        __syncthreads();
        // trailing update: the wavefronts take the rows in turn, the lanes the
        // columns (no index division; every entry sees the same single fma)
        for (int i = c + 1 + (tid >> 6); i < M; i += MIX_THREADS / 64) {
            const float l = Af[i * LDA + c];
            for (int j = c + 1 + (tid & 63); j < M; j += 64)
                Af[i * LDA + j] = fmaf(-l, Af[c * LDA + j], Af[i * LDA + j]);
        }
        __syncthreads();
    }
    // W (FP32) holds the current right-hand side; X accumulates in FP64
    for (int t = tid; t < M * 4; t += blockDim.x) {
        const int i = t >> 2, q = t & 3;
        W[t] = (float)B[perm[i] * 4 + q];
        X[t] = 0.0;
    }
    __syncthreads();
    int converged = 0;
    double dprev = 1e300;
    for (int it = 0; it <= max_ref && !singular; ++it) {
        // forward (unit L) and backward (U) substitution on the nr columns of W
        // (four columns per row, t = 4 i + q; column q >= nr is not used)
        for (int c = 0; c < M; ++c) {
            for (int t = tid; t < (M - c - 1) * 4; t += blockDim.x) {
                const int i = c + 1 + (t >> 2), q = t & 3;
                if (q < nr) W[i * 4 + q] = fmaf(-Af[i * LDA + c], W[c * 4 + q], W[i * 4 + q]);
            }
            __syncthreads();
        }
        for (int c = M - 1; c >= 0; --c) {
            if (tid < nr) W[c * 4 + tid] = W[c * 4 + tid] / Af[c * LDA + c];
            __syncthreads();
            for (int t = tid; t < c * 4; t += blockDim.x) {
                const int i = t >> 2, q = t & 3;
                if (q < nr) W[i * 4 + q] = fmaf(-Af[i * LDA + c], W[c * 4 + q], W[i * 4 + q]);
            }
            __syncthreads();
        }
        // x += dx; size of the correction relative to x
        if (tid == 0) {
            s_red[0] = 0.0;
            s_red[1] = 0.0;
        }
        __syncthreads();
        double dmax = 0.0, xmax = 0.0;
        for (int t = tid; t < M * 4; t += blockDim.x) {
            if ((t & 3) >= nr) continue;
            const double dx = (double)W[t];
            const double xn = X[t] + dx;
            X[t] = xn;
            dmax = fmax(dmax, fabs(dx));
            xmax = fmax(xmax, fabs(xn));
        }
        for (int off = 32; off > 0; off >>= 1) {
            dmax = fmax(dmax, __shfl_xor(dmax, off));
            xmax = fmax(xmax, __shfl_xor(xmax, off));
        }
        if ((tid & 63) == 0) {
            atomicMax((unsigned long long*)&s_red[0], (unsigned long long)__double_as_longlong(dmax));
            atomicMax((unsigned long long*)&s_red[1], (unsigned long long)__double_as_longlong(xmax));
        }
        __syncthreads();
        // converged: the correction is at the FP64 rounding floor (<= 1e-13 of x),
        // or it has stopped shrinking below 1e-11 of x (residual noise ~ cond * eps64)
        const double dcur = s_red[0], xcur = s_red[1];
        if (dcur <= 1e-13 * xcur || (it > 0 && dcur <= 1e-11 * xcur && dcur >= 0.5 * dprev)) {
            converged = 1;
            break;
        }
        if (it == max_ref || !(dcur == dcur)) break;
        dprev = dcur;
        // r = b - A x in FP64 (entries recomputed), permuted into W as FP32
        for (int i = tid; i < M; i += blockDim.x) {
            const int pi = perm[i];
            double acc[4] = {B[pi * 4 + 0], B[pi * 4 + 1], B[pi * 4 + 2], B[pi * 4 + 3]};
            for (int j = 0; j < M; ++j) {
                const double a = saddle_entry(pi, j, k, dim, phs, xs, s_e);
                if (a != 0.0)
                    for (int q = 0; q < 4; ++q) acc[q] = fma(-a, X[j * 4 + q], acc[q]);
            }
            for (int q = 0; q < 4; ++q) W[i * 4 + q] = (float)acc[q];
        }
        __syncthreads();
    }
    if (tid == 0) s_conv = converged;
    __syncthreads();
    if (!s_conv) {
        if (tid == 0) flag[r] = 1;
        return;
    }
    for (int t = tid; t < k; t += blockDim.x) {
        const size_t slot = (size_t)base + (size_t)sell_off(t);
        gx[slot] = X[t * 4 + 0];
        gy[slot] = X[t * 4 + 1];
        if (dim == 3) gz[slot] = X[t * 4 + 2];
        lap[slot] = X[t * 4 + dim];
    }
}

}  // namespace

size_t weights_mixed_lds(int max_len, int m) {
    const int M = max_len + m;
    return sizeof(double) * ((size_t)M * 8 + (size_t)max_len * 3) +
           sizeof(float) * ((size_t)M * (M + 1) + (size_t)M * 4) + sizeof(int) * M;
}

void launch_weights_mixed(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                          const SellView& sv, const double* xyz_soa, int npad, double* gx,
                          double* gy, double* gz, double* lap, const int* rows, int nrows,
                          int* flag, int max_len, int max_ref) {
    if (nrows == 0) return;
    if (m > WAVE_MAX_POLY || weights_mixed_lds(max_len, m) > WEIGHTS_LDS_MAX)
        throw std::runtime_error("launch_weights_mixed: system too large for the mixed kernel");
    const size_t lds = weights_mixed_lds(max_len, m);
    if (lds > 64 * 1024)
        MP_HIP(hipFuncSetAttribute((const void*)weights_mixed_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(weights_mixed_kernel, dim3(nrows), dim3(MIX_THREADS), lds, s, dim, phs, m, exps_dev,
                       sv, xyz_soa, npad, gx, gy, gz, lap, rows, flag, max_ref);
    MP_HIP(hipGetLastError());
}

void launch_weights_fp64(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                         const SellView& sv, const double* xyz_soa, int npad, double* gx,
                         double* gy, double* gz, double* lap, double* cond, int max_len,
                         const int* rows, int nrows, const int* only) {
    if (rows && nrows == 0) return;
    const int n = rows ? nrows : sv.nv;
    const int M = max_len + m;
    const size_t doubles = (size_t)M * (M + 1) + (size_t)M * 4 + (size_t)max_len * 3;
    const size_t lds = sizeof(double) * doubles;
    if (lds <= WEIGHTS_LDS_MAX) {
        if (lds > 64 * 1024)
            MP_HIP(hipFuncSetAttribute((const void*)weights_fp64_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(weights_fp64_kernel, dim3(n), dim3(256), lds, s, dim, phs, m, exps_dev,
                           sv, xyz_soa, npad, gx, gy, gz, lap, cond, rows, only, 0,
                           (double*)nullptr, 0L);
        MP_HIP(hipGetLastError());
        return;
    }
    // systems too large for LDS (high polynomial degrees): a global work area
    // of `doubles` per block, in launches of at most 256 MiB of it
    const long stride = (long)((doubles + 31) & ~(size_t)31);
    const int chunk = (int)std::max<long>(1, std::min<long>(n, (256L << 20) / 8 / stride));
    DevBuf<double> work;
    work.alloc((size_t)chunk * (size_t)stride);
    for (int t0 = 0; t0 < n; t0 += chunk) {
        const int cnt = std::min(chunk, n - t0);
        hipLaunchKernelGGL(weights_fp64_kernel, dim3(cnt), dim3(256), 0, s, dim, phs, m, exps_dev,
                           sv, xyz_soa, npad, gx, gy, gz, lap, cond, rows, only, t0, work.p,
                           stride);
        MP_HIP(hipGetLastError());
    }
    MP_HIP(hipStreamSynchronize(s));  // (the work area is freed on return)
}


template <int NW, int AW>
void launch_wave(hipStream_t s, int dim, int phs, int m, const int* exps_dev, const SellView& sv,
                 const double* xyz_soa, int npad, double* gx, double* gy, double* gz, double* lap,
                 const int* rows, int nrows, int* redo) {
    if (nrows == 0) return;
    const size_t lds = sizeof(double) * (size_t)wave_lds_doubles(AW, AW - m, m);
    if (lds > 64 * 1024)
        MP_HIP(hipFuncSetAttribute((const void*)weights_wave_kernel<NW, AW>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((weights_wave_kernel<NW, AW>), dim3(nrows), dim3(64 * NW), lds, s, dim,
                       phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, rows, redo);
    MP_HIP(hipGetLastError());
}

void launch_weights_fp64_rows(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                              const SellView& sv, const double* xyz_soa, int npad, double* gx,
                              double* gy, double* gz, double* lap, const std::vector<int>& rows,
                              const std::vector<int>& lens, DevBuf<int>& dev_rows) {
    static const int cls[WAVE_CLASSES] = {32, 48, 64, 80, 96};
    std::vector<int> bucket[WAVE_CLASSES + 1];
    int max_long = 0;
    for (size_t t = 0; t < rows.size(); ++t) {
        const int M = lens[t] + m;
        int b = 0;
        while (b < WAVE_CLASSES && M > cls[b]) ++b;
        if (m > WAVE_MAX_POLY) b = WAVE_CLASSES;  // (the register kernels' monomial table)
        bucket[b].push_back(rows[t]);
        if (b == WAVE_CLASSES) max_long = std::max(max_long, lens[t]);
    }
    std::vector<int> all;
    all.reserve(rows.size());
    int off[WAVE_CLASSES + 2];
    for (int b = 0; b <= WAVE_CLASSES; ++b) {
        off[b] = (int)all.size();
        all.insert(all.end(), bucket[b].begin(), bucket[b].end());
    }
    off[WAVE_CLASSES + 1] = (int)all.size();
    if (all.empty()) return;
    // the grouped row list, then one flag per row (set by a single-wavefront
    // stencil that met an exactly zero pivot)
    const size_t nall = all.size();
    dev_rows.alloc(2 * nall);
    MP_HIP(hipMemcpyAsync(dev_rows.p, all.data(), sizeof(int) * nall, hipMemcpyHostToDevice, s));
    MP_HIP(hipMemsetAsync(dev_rows.p + nall, 0, sizeof(int) * nall, s));
    auto n = [&](int b) { return off[b + 1] - off[b]; };
    const int* d = dev_rows.p;
    int* redo = dev_rows.p + nall;
    launch_wave<1, 32>(s, dim, phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, d + off[0],
                       n(0), redo + off[0]);
    launch_wave<1, 48>(s, dim, phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, d + off[1],
                       n(1), redo + off[1]);
    launch_wave<1, 64>(s, dim, phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, d + off[2],
                       n(2), redo + off[2]);
    launch_wave<2, 80>(s, dim, phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, d + off[3],
                       n(3), redo + off[3]);
    launch_wave<2, 96>(s, dim, phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, d + off[4],
                       n(4), redo + off[4]);
    // single-wavefront stencils with a zero pivot: the LDS kernel, which
    // follows the oracle through the singular column
    const int n1 = off[3] - off[0];
    if (n1 > 0)
        launch_weights_fp64(s, dim, phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, nullptr,
                            64 - m, d + off[0], n1, redo + off[0]);
    if (n(WAVE_CLASSES) > 0)
        launch_weights_fp64(s, dim, phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, nullptr,
                            max_long, d + off[WAVE_CLASSES], n(WAVE_CLASSES));
}

}  // namespace mp
