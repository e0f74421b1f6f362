// Initial guess of the pressure solve by projection onto the last K solutions:
// the combination of the stored pressures whose residual against the new
// right-hand side is smallest (P. F. Fischer, "Projection techniques for
// iterative solution of Ax = b with successive right-hand sides", Comput.
// Methods Appl. Mech. Engrg. 163 (1998) 193-204).  The stored pressures p_j
// (newest first) and their products A p_j live in rings of PROJ_MAX slots.
// Successive pressures are nearly parallel, so the least-squares problem is
// posed in the basis of backward differences in time, d_k = (Delta^k p) at the
// newest step, each column scaled to unit norm; its normal equations are then
// well conditioned enough for FP64 Cholesky:
//   G_kl = (A d_k).(A d_l),  c_k = (A d_k).b,  (S G S + eps I) y = S c,
//   beta = S y,  x0 = sum_k beta_k d_k   (S = diag(1/|A d_k|)).
// The differences of A p are formed per row from the stored products (no
// product with A at guess time); A p of each new solution comes from the
// solve's own final FP64 residual, A p = b - (b - A p).
#include "mp_internal.hpp"

namespace mp {

namespace {

constexpr int PROJ_THREADS = 256;
constexpr int PROJ_BLOCKS = 256;

__device__ __forceinline__ int ring_slot(int head, int j) {
    return (head - j + PROJ_MAX) % PROJ_MAX;
}

// d[k] = Delta^k v at the newest entry (Newton's backward differences of
// v[0] (newest) .. v[K-1])
template <int K>
__device__ __forceinline__ void backward_differences(double (&v)[K], double (&d)[K]) {
    d[0] = v[0];
#pragma unroll
    for (int k = 1; k < K; ++k) {
#pragma unroll
        for (int j = 0; j < K - k; ++j) v[j] = v[j] - v[j + 1];
        d[k] = v[0];
    }
}

// Gram entry q of the K differences (q < K(K+1)/2: the pair (k, l), k <= l,
// in row order of the upper triangle; after them c_k = d_k . b): its two
// factor indices, K standing for b.
constexpr int gram_a(int K, int q) {
    for (int k = 0; k < K; ++k) {
        if (q < K - k) return k;
        q -= K - k;
    }
    return q;
}
constexpr int gram_b(int K, int q) {
    for (int k = 0; k < K; ++k) {
        if (q < K - k) return k + q;
        q -= K - k;
    }
    return K;
}
// threads per row: windows above 12 split each row's Gram entries over two
// threads (the second loads the row again, from cache), so the accumulators
// stay in registers (beyond 16 the compiler no longer keeps the unrolled
// loops in registers)
template <int K>
constexpr int gram_split() { return K <= 12 ? 1 : 2; }
static_assert(PROJ_MAX <= 16, "windows up to 16 stay in registers");

template <int K>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_gram(int n, const double* __restrict__ hw,
                                                            size_t ld, const ProjState* ps,
                                                            const double* __restrict__ b,
                                                            double* __restrict__ partial) {
    constexpr int NG = K * (K + 1) / 2 + K;
    constexpr int NS = gram_split<K>();
    constexpr int NPER = (NG + NS - 1) / NS;
    double acc[NPER];
#pragma unroll
    for (int j = 0; j < NPER; ++j) acc[j] = 0.0;
    const int head = ps->head;
    const int part = NS == 1 ? 0 : (int)(threadIdx.x % NS);
    for (int i = (blockIdx.x * PROJ_THREADS + threadIdx.x) / NS; i < n;
         i += gridDim.x * (PROJ_THREADS / NS)) {
        double w[K], d[K];
#pragma unroll
        for (int j = 0; j < K; ++j) w[j] = hw[(size_t)ring_slot(head, j) * ld + i];
        backward_differences<K>(w, d);
        const double bi = b[i];
        // (unrolled: every index below is a constant, so d stays in registers)
#pragma unroll
        for (int j = 0; j < NPER; ++j) {
            // this thread's entry j: part p takes entry p * NPER + j
            double u = 0.0, v = 0.0;
            bool any = false;
#pragma unroll
            for (int p = 0; p < NS; ++p) {
                const int q = p * NPER + j;
                if (q < NG) {
                    const int bq = gram_b(K, q);
                    const double uq = d[gram_a(K, q)], vq = bq >= K ? bi : d[bq];
                    u = part == p ? uq : u;
                    v = part == p ? vq : v;
                    any = any || part == p;
                }
            }
            if (any) acc[j] = fma(u, v, acc[j]);
        }
    }
    // block sums per entry: the shuffle tree keeps the row parts apart
    __shared__ double sh[PROJ_THREADS / 64][NG];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NPER; ++j) {
        double v = acc[j];
        for (int off = 32; off >= NS; off >>= 1) v += __shfl_xor(v, off);
        const int q = lane * NPER + j;  // lane < NS holds part `lane`
        if (lane < NS && q < NG) sh[wid][q] = v;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NG; q += PROJ_THREADS) {
        double v = sh[0][q];
        for (int u = 1; u < PROJ_THREADS / 64; ++u) v += sh[u][q];
        partial[(size_t)blockIdx.x * NG + q] = v;
    }
}

// Sums the block partials (wavefront u sums entries u, u+4, ...: each lane
// the partials of blocks lane, lane+64, ..., then a shuffle tree), then one
// wavefront solves for beta: lane r holds row r of the scaled Gram matrix,
// and the right-looking Cholesky and the two triangular solves pass columns
// between lanes by shuffles.
template <int K>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_solve(const double* __restrict__ partial,
                                                             ProjState* ps) {
    constexpr int NG = K * (K + 1) / 2 + K;
    static_assert(PROJ_BLOCKS % 64 == 0, "whole wavefronts of block partials");
    static_assert(K <= 64, "one row per lane");
    __shared__ double tot[NG];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll 8
    for (int q = wid; q < NG; q += PROJ_THREADS / 64) {
        double a = 0.0;
#pragma unroll
        for (int blk = lane; blk < PROJ_BLOCKS; blk += 64) a += partial[(size_t)blk * NG + q];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
        if (lane == 0) tot[q] = a;
    }
    __syncthreads();
    if (wid != 0) return;
    const int r = lane < K ? lane : 0;  // (lanes past K mirror row 0 and are not used)
    auto idx = [](int k, int l) { return k * K - k * (k - 1) / 2 + (l - k); };  // k <= l
    double g[K];
#pragma unroll
    for (int l = 0; l < K; ++l) g[l] = tot[r <= l ? idx(r, l) : idx(l, r)];
    const double grr = tot[idx(r, r)];
    // unit columns; a column with no length (no change in time) drops out
    const double s = grr > 0.0 ? 1.0 / sqrt(grr) : 0.0;
    double c = tot[K * (K + 1) / 2 + r] * s;
#pragma unroll
    for (int l = 0; l < K; ++l) {
        g[l] *= s * __shfl(s, l);
        if (l == r) g[l] += 1e-13;  // (keeps the factorisation defined for dependent columns)
    }
    // Cholesky, column by column: lane r's g[k] becomes L[r][k] (r >= k)
    bool ok = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double dkk = __shfl(g[k], k);
        ok = ok && dkk > 0.0;
        const double lkk = sqrt(ok ? dkk : 1.0);
        g[k] = r > k ? g[k] / lkk : (r == k ? lkk : g[k]);
#pragma unroll
        for (int m = k + 1; m < K; ++m) {
            const double lmk = __shfl(g[k], m);
            if (r > k) g[m] -= g[k] * lmk;
        }
    }
    // L y = c, then L^T x = y (y, x the same in every lane)
    double y[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        y[k] = __shfl(c, k) / __shfl(g[k], k);
        if (r > k) c -= g[k] * y[k];
    }
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        y[k] = y[k] / __shfl(g[k], k);
#pragma unroll
        for (int m = 0; m < k; ++m) y[m] -= __shfl(g[m], k) * y[k];
    }
    double sk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sk[k] = __shfl(s, k);
    if (lane != 0) return;
#pragma unroll
    for (int k = 0; k < PROJ_MAX; ++k) {
        double bk = 0.0;
        if (k < K) bk = ok ? (isfinite(y[k]) ? y[k] : 0.0) * sk[k] : (k == 0 ? 1.0 : 0.0);
        ps->beta[k] = bk;  // (no factorisation: the newest pressure alone)
    }
}

template <int K>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_combine(int n, const double* __restrict__ hp,
                                                               size_t ld, const ProjState* ps,
                                                               double* __restrict__ out) {
    const int i = blockIdx.x * PROJ_THREADS + threadIdx.x;
    if (i >= n) return;
    const int head = ps->head;
    double v[K], d[K];
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] = hp[(size_t)ring_slot(head, j) * ld + i];
    backward_differences<K>(v, d);
    double x = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) x = fma(ps->beta[k], d[k], x);
    out[i] = x;
}

// The new solution and its product, A p = b - r, into the slot after head.
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_store(int n, const double* __restrict__ p,
                                                             const double* __restrict__ b,
                                                             const double* __restrict__ r,
                                                             double* __restrict__ hp,
                                                             double* __restrict__ hw, size_t ld,
                                                             const ProjState* ps) {
    const int i = blockIdx.x * PROJ_THREADS + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)((ps->head + 1) % PROJ_MAX) * ld + i;
    hp[o] = p[i];
    hw[o] = b[i] - r[i];
}

__global__ void k_proj_advance(ProjState* ps, int reset) {
    if (reset) {
        ps->head = 0;
        ps->count = 0;
        return;
    }
    ps->head = (ps->head + 1) % PROJ_MAX;
    ps->count = ps->count < PROJ_MAX ? ps->count + 1 : PROJ_MAX;
}

template <int K>
void launch_guess_k(hipStream_t s, int n, const double* hp, const double* hw, size_t ld,
                    ProjState* ps, const double* b, double* partial, double* out) {
    hipLaunchKernelGGL(k_proj_gram<K>, dim3(PROJ_BLOCKS), dim3(PROJ_THREADS), 0, s, n, hw, ld, ps,
                       b, partial);
    hipLaunchKernelGGL(k_proj_solve<K>, dim3(1), dim3(PROJ_THREADS), 0, s, partial, ps);
    hipLaunchKernelGGL(k_proj_combine<K>, dim3((n + PROJ_THREADS - 1) / PROJ_THREADS),
                       dim3(PROJ_THREADS), 0, s, n, hp, ld, ps, out);
}

}  // namespace

size_t proj_partial_size() { return (size_t)PROJ_BLOCKS * (PROJ_MAX * (PROJ_MAX + 1) / 2 + PROJ_MAX); }

void launch_proj_guess(hipStream_t s, int k, int n, const double* hp, const double* hw, size_t ld,
                       ProjState* ps, const double* b, double* partial, double* out) {
    switch (k) {
        case 1: launch_guess_k<1>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 2: launch_guess_k<2>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 3: launch_guess_k<3>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 4: launch_guess_k<4>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 5: launch_guess_k<5>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 6: launch_guess_k<6>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 7: launch_guess_k<7>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 8: launch_guess_k<8>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 9: launch_guess_k<9>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 10: launch_guess_k<10>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 11: launch_guess_k<11>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 12: launch_guess_k<12>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 13: launch_guess_k<13>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 14: launch_guess_k<14>(s, n, hp, hw, ld, ps, b, partial, out); break;
        case 15: launch_guess_k<15>(s, n, hp, hw, ld, ps, b, partial, out); break;
        default: launch_guess_k<16>(s, n, hp, hw, ld, ps, b, partial, out); break;
    }
    MP_HIP(hipGetLastError());
}

void launch_proj_store(hipStream_t s, int n, const double* p, const double* b, const double* r,
                       double* hp, double* hw, size_t ld, ProjState* ps) {
    hipLaunchKernelGGL(k_proj_store, dim3((n + PROJ_THREADS - 1) / PROJ_THREADS),
                       dim3(PROJ_THREADS), 0, s, n, p, b, r, hp, hw, ld, ps);
    hipLaunchKernelGGL(k_proj_advance, dim3(1), dim3(1), 0, s, ps, 0);
    MP_HIP(hipGetLastError());
}

void launch_proj_reset(hipStream_t s, ProjState* ps) {
    hipLaunchKernelGGL(k_proj_advance, dim3(1), dim3(1), 0, s, ps, 1);
    MP_HIP(hipGetLastError());
}

}  // namespace mp
