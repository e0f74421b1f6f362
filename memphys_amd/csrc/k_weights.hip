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

__global__ __launch_bounds__(256) void weights_fp64_kernel(
    int dim, int phs, int m, const int* __restrict__ exps, SellView sv,
    const double* __restrict__ xyz, int npad, double* __restrict__ gx, double* __restrict__ gy,
    double* __restrict__ gz, double* __restrict__ lap, double* __restrict__ cond,
    const int* __restrict__ rows) {
    extern __shared__ double lds[];
    __shared__ double s_mn[3], s_sc[3];
    __shared__ double s_anorm, s_red[4];
    __shared__ int s_piv, s_jm;
    __shared__ int s_e[MAX_POLY * 3];

    const int r = rows ? rows[blockIdx.x] : (int)blockIdx.x;
    if (r >= sv.nv || sv.len[r] == 0) return;  // padding rows carry no stencil
    const int tid = threadIdx.x;
    const int k = sv.len[r];
    const int M = k + m;
    const int LDA = M + 1;
    const int nr = dim + 1;
    double* A = lds;
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
    const double p = (double)phs;
    for (int i = tid; i < M; i += blockDim.x) {
        if (i < k) {
            double r2 = 0.0;
            for (int a = 0; a < dim; ++a) {
                double d = xs[i * 3 + a] - xs[a];
                r2 = r2 + d * d;
            }
            for (int a = 0; a < dim; ++a) {
                double v = (-p) * (xs[i * 3 + a] - xs[a]);
                v = v * rpow_d(r2, phs - 2);
                R[i * 4 + a] = v / s_sc[a];
            }
            double lp = 0.0;
            if (fabs(r2) > 1e-14) {
                double g[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < dim; ++a) {
                    double d = xs[i * 3 + a] - xs[a];
                    double d2 = d * d;
                    double q4 = rpow_d(r2, phs - 4);
                    double v = (p * p) * d2 * q4;
                    v = v - (2.0 * p) * d2 * q4;
                    v = v + p * rpow_d(r2, phs - 2);
                    g[a] = v / (s_sc[a] * s_sc[a]);
                }
                lp = g[0] + g[1];
                if (dim == 3) lp = lp + g[2];
            }
            R[i * 4 + dim] = lp;
        } else {
            const int t = i - k;
            const int* e = s_e + t * dim;
            const double x0 = xs[0], x1 = xs[1], x2 = (dim == 3) ? xs[2] : 0.0;
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
                    v = v / s_sc[a];
                }
                R[i * 4 + a] = v;
            }
            double g0 = 0.0, g1 = 0.0, g2 = 0.0;
            if (e[0] > 1) {
                double v = (double)e[0] * ((double)e[0] - 1.0) * ipow_d(x0, e[0] - 2);
                v = v * ipow_d(x1, e[1]);
                if (dim == 3) v = v * ipow_d(x2, e[2]);
                g0 = v / (s_sc[0] * s_sc[0]);
            }
            if (e[1] > 1) {
                double v = (double)e[1] * ((double)e[1] - 1.0) * ipow_d(x1, e[1] - 2);
                v = v * ipow_d(x0, e[0]);
                if (dim == 3) v = v * ipow_d(x2, e[2]);
                g1 = v / (s_sc[1] * s_sc[1]);
            }
            if (dim == 3 && e[2] > 1) {
                double v = (double)e[2] * ((double)e[2] - 1.0) * ipow_d(x2, e[2] - 2);
                v = v * ipow_d(x0, e[0]);
                v = v * ipow_d(x1, e[1]);
                g2 = v / (s_sc[2] * s_sc[2]);
            }
            double lp = g0 + g1;
            if (dim == 3) lp = lp + g2;
            R[i * 4 + dim] = lp;
        }
    }
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
__global__ __launch_bounds__(256) void weights_mixed_kernel(
    int dim, int phs, int m, const int* __restrict__ exps, SellView sv,
    const double* __restrict__ xyz, int npad, double* __restrict__ gx, double* __restrict__ gy,
    double* __restrict__ gz, double* __restrict__ lap, const int* __restrict__ rows,
    int* __restrict__ flag, int max_ref) {
    extern __shared__ double lds[];
    __shared__ double s_mn[3], s_sc[3];
    __shared__ float s_bestf;
    __shared__ int s_piv, s_conv;
    __shared__ double s_red[2];
    __shared__ int s_e[MAX_POLY * 3];
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
    for (int idx = tid; idx < M * M; idx += blockDim.x) {
        const int i = idx / M, j = idx - i * M;
        Af[i * LDA + j] = (float)saddle_entry(i, j, k, dim, phs, xs, s_e);
    }
    // FP64 right-hand sides (same formulas as the FP64 kernel)
    const double p = (double)phs;
    for (int i = tid; i < M; i += blockDim.x) {
        for (int q = 0; q < 4; ++q) B[i * 4 + q] = 0.0;
        if (i < k) {
            double r2 = 0.0;
            for (int a = 0; a < dim; ++a) {
                double d = xs[i * 3 + a] - xs[a];
                r2 = r2 + d * d;
            }
            for (int a = 0; a < dim; ++a) {
                double v = (-p) * (xs[i * 3 + a] - xs[a]);
                v = v * rpow_d(r2, phs - 2);
                B[i * 4 + a] = v / s_sc[a];
            }
            double lp = 0.0;
            if (fabs(r2) > 1e-14) {
                double g[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < dim; ++a) {
                    double d = xs[i * 3 + a] - xs[a];
                    double d2 = d * d;
                    double q4 = rpow_d(r2, phs - 4);
                    double v = (p * p) * d2 * q4;
                    v = v - (2.0 * p) * d2 * q4;
                    v = v + p * rpow_d(r2, phs - 2);
                    g[a] = v / (s_sc[a] * s_sc[a]);
                }
                lp = g[0] + g[1];
                if (dim == 3) lp = lp + g[2];
            }
            B[i * 4 + dim] = lp;
        } else {
            const int t = i - k;
            const int* e = s_e + t * dim;
            const double x0 = xs[0], x1 = xs[1], x2 = (dim == 3) ? xs[2] : 0.0;
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
                    v = v / s_sc[a];
                }
                B[i * 4 + a] = v;
            }
            double g0 = 0.0, g1 = 0.0, g2 = 0.0;
            if (e[0] > 1) {
                double v = (double)e[0] * ((double)e[0] - 1.0) * ipow_d(x0, e[0] - 2);
                v = v * ipow_d(x1, e[1]);
                if (dim == 3) v = v * ipow_d(x2, e[2]);
                g0 = v / (s_sc[0] * s_sc[0]);
            }
            if (e[1] > 1) {
                double v = (double)e[1] * ((double)e[1] - 1.0) * ipow_d(x1, e[1] - 2);
                v = v * ipow_d(x0, e[0]);
                if (dim == 3) v = v * ipow_d(x2, e[2]);
                g1 = v / (s_sc[1] * s_sc[1]);
            }
            if (dim == 3 && e[2] > 1) {
                double v = (double)e[2] * ((double)e[2] - 1.0) * ipow_d(x2, e[2] - 2);
                v = v * ipow_d(x0, e[0]);
                v = v * ipow_d(x1, e[1]);
                g2 = v / (s_sc[2] * s_sc[2]);
            }
            double lp = g0 + g1;
            if (dim == 3) lp = lp + g2;
            B[i * 4 + dim] = lp;
        }
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
        const int nrow = M - c - 1;
        for (int idx = tid; idx < nrow * nrow; idx += blockDim.x) {
            const int ii = idx / nrow;
            const int i = c + 1 + ii, j = c + 1 + (idx - ii * nrow);
            Af[i * LDA + j] = fmaf(-Af[i * LDA + c], Af[c * LDA + j], Af[i * LDA + j]);
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
        for (int c = 0; c < M; ++c) {
            for (int t = tid; t < (M - c - 1) * nr; t += blockDim.x) {
                const int i = c + 1 + t / nr, q = t % nr;
                W[i * 4 + q] = fmaf(-Af[i * LDA + c], W[c * 4 + q], W[i * 4 + q]);
            }
            __syncthreads();
        }
        for (int c = M - 1; c >= 0; --c) {
            if (tid < nr) W[c * 4 + tid] = W[c * 4 + tid] / Af[c * LDA + c];
            __syncthreads();
            for (int t = tid; t < c * nr; t += blockDim.x) {
                const int i = t / nr, q = t % nr;
                W[i * 4 + q] = fmaf(-Af[i * LDA + c], W[c * 4 + q], W[i * 4 + q]);
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

void launch_weights_mixed(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                          const SellView& sv, const double* xyz_soa, int npad, double* gx,
                          double* gy, double* gz, double* lap, const int* rows, int nrows,
                          int* flag, int max_len, int max_ref) {
    if (nrows == 0) return;
    const int M = max_len + m;
    const size_t lds = sizeof(double) * ((size_t)M * 8 + (size_t)max_len * 3) +
                       sizeof(float) * ((size_t)M * (M + 1) + (size_t)M * 4) + sizeof(int) * M;
    if (lds > 64 * 1024)
        MP_HIP(hipFuncSetAttribute((const void*)weights_mixed_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(weights_mixed_kernel, dim3(nrows), dim3(256), lds, s, dim, phs, m, exps_dev,
                       sv, xyz_soa, npad, gx, gy, gz, lap, rows, flag, max_ref);
    MP_HIP(hipGetLastError());
}

void launch_weights_fp64(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                         const SellView& sv, const double* xyz_soa, int npad, double* gx,
                         double* gy, double* gz, double* lap, double* cond, int max_len,
                         const int* rows, int nrows) {
    if (rows && nrows == 0) return;
    const int M = max_len + m;
    const size_t lds = sizeof(double) * ((size_t)M * (M + 1) + (size_t)M * 4 + (size_t)max_len * 3);
    if (lds > 64 * 1024)
        MP_HIP(hipFuncSetAttribute((const void*)weights_fp64_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(weights_fp64_kernel, dim3(rows ? nrows : sv.nv), dim3(256), lds, s, dim,
                       phs, m, exps_dev, sv, xyz_soa, npad, gx, gy, gz, lap, cond, rows);
    MP_HIP(hipGetLastError());
}

}  // namespace mp
