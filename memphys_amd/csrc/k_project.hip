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
constexpr int PROJ_BLOCKS = 768;  // three per CU (LDS holds three up to K = 20)
constexpr int PROJ_TILE = PROJ_THREADS;  // rows per pass of a Gram block, one per thread
static_assert(PROJ_MAX < 32, "the Gram matrix with b fits two 16-column blocks");

typedef double d4 __attribute__((ext_vector_type(4)));

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

// The Gram pass on the matrix cores.  The tile's rows hold the K differences
// and b (columns 0..K-1 and K); v_mfma_f64_16x16x4f64 adds, per four rows,
// C[i][j] += sum_r D[r][16 I + i] D[r][16 J + j] over 16-column blocks I, J
// (lane l supplies D[r = l/16][16 I + l%16] as both operands' element, and
// holds C[l/16 + 4 v][l % 16] in its register v).  Tiles: K < 16, the one
// block (b in column K); K = 16, block 0 against itself and against b alone;
// K > 16, blocks (0,0), (0,1) and (1,1), b in column K of block 1.
template <int K>
constexpr int gram_tiles() { return K < 16 ? 1 : (K == 16 ? 2 : 3); }
// Row stride of the tile in LDS (doubles): odd, so a quarter-wavefront's
// sixteen consecutive columns and a wavefront's rows spread over the banks.
template <int K>
constexpr int tile_ld() { return (K + 1) % 2 ? K + 1 : K + 2; }
template <int K>
constexpr int gram_lds() {
    constexpr int a = PROJ_TILE * tile_ld<K>() + 32;  // (+32: the second block's reads past the last row)
    constexpr int b = (PROJ_THREADS / 64) * gram_tiles<K>() * 64 * 4;
    return a > b ? a : b;
}

// Gram entry (k, l), k <= l, l = K standing for b: where a wavefront's
// accumulators hold it (tile, lane, register).
template <int K>
__device__ __forceinline__ int gram_slot(int k, int l) {
    int t, i, j;
    if (K < 16) {
        t = 0, i = k, j = l;
    } else if (K == 16) {
        t = l == K ? 1 : 0, i = k, j = l == K ? 0 : l;
    } else {
        t = k >= 16 ? 2 : (l >= 16 ? 1 : 0), i = k & 15, j = l & 15;
    }
    return (t * 64 + j + 16 * (i & 3)) * 4 + (i >> 2);
}

template <int K>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_gram(int n, const double* __restrict__ hw,
                                                            size_t ld, const ProjState* ps,
                                                            const double* __restrict__ b,
                                                            double* __restrict__ partial) {
    constexpr int NG = K * (K + 1) / 2 + K;
    constexpr int NT = gram_tiles<K>();
    constexpr int LDT = tile_ld<K>();
    __shared__ double sd[gram_lds<K>()];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int col = lane & 15, sub = lane >> 4;
    const int head = ps->head;
    // two accumulator sets (even and odd row groups) so consecutive matrix
    // operations do not wait on each other
    const d4 zero = {0.0, 0.0, 0.0, 0.0};
    d4 acc[2][NT];
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[e][t] = zero;
    const int ntiles = (n + PROJ_TILE - 1) / PROJ_TILE;
    // the next tile's row is loaded while the matrix cores work on this one
    double w[K], bi = 0.0;
    auto load = [&](int t) {
        const int i = t * PROJ_TILE + threadIdx.x;
#pragma unroll
        for (int j = 0; j < K; ++j) w[j] = i < n ? hw[(size_t)ring_slot(head, j) * ld + i] : 0.0;
        bi = i < n ? b[i] : 0.0;
    };
    if (blockIdx.x < ntiles) load(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        double* row = sd + threadIdx.x * LDT;
        double d[K];
        backward_differences<K>(w, d);  // (rows past n hold zeros)
#pragma unroll
        for (int k = 0; k < K; ++k) row[k] = d[k];
        row[K] = bi;
        __syncthreads();
        if (tile + (int)gridDim.x < ntiles) load(tile + gridDim.x);
        // wavefront wid takes the four-row groups wid, wid + 4, ..., alternately
        // into its two accumulator sets
        constexpr int NW = PROJ_THREADS / 64;
        static_assert((PROJ_TILE / 4) % (2 * NW) == 0, "whole pairs of row groups per wavefront");
#pragma unroll 2
        for (int it = 0; it < PROJ_TILE / 4 / (2 * NW); ++it) {
            const int g0 = wid + 2 * NW * it;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const double* rr = sd + (4 * (g0 + e * NW) + sub) * LDT;
                const double a0 = col <= K ? rr[col] : 0.0;
                if constexpr (K < 16) {
                    acc[e][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc[e][0], 0, 0, 0);
                } else if constexpr (K == 16) {
                    const double bb = col == 0 ? rr[K] : 0.0;
                    acc[e][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc[e][0], 0, 0, 0);
                    acc[e][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bb, acc[e][1], 0, 0, 0);
                } else {
                    const double a1 = 16 + col <= K ? rr[16 + col] : 0.0;
                    acc[e][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc[e][0], 0, 0, 0);
                    acc[e][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a1, acc[e][1], 0, 0, 0);
                    acc[e][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, acc[e][2], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // block sums: each wavefront's accumulators through LDS, then one thread per entry
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const d4 s = acc[0][t] + acc[1][t];
        double* dst = sd + ((wid * NT + t) * 64 + lane) * 4;
        dst[0] = s[0];
        dst[1] = s[1];
        dst[2] = s[2];
        dst[3] = s[3];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NG; q += PROJ_THREADS) {
        int k, l;
        if (q < K * (K + 1) / 2) {
            int r = q;  // (row k of the upper triangle holds K - k entries)
            k = 0;
            while (r >= K - k) {
                r -= K - k;
                ++k;
            }
            l = k + r;
        } else {
            k = q - K * (K + 1) / 2;
            l = K;
        }
        const int s = gram_slot<K>(k, l);
        double v = 0.0;
#pragma unroll
        for (int u = 0; u < PROJ_THREADS / 64; ++u) v += sd[u * NT * 64 * 4 + s];
        partial[(size_t)blockIdx.x * NG + q] = v;
    }
}

// Value of lane src (wave-uniform) in every lane.
__device__ __forceinline__ double lane_value(double v, int src) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sums the Gram blocks' partial sums: block q adds entry q over the blocks
// (a fixed order, so the sums repeat bit for bit).  One block per entry keeps
// the many loads in flight; a single block walking every entry waited on them
// (110 us for K = 20 against a few here).
template <int K>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_reduce(const double* __restrict__ partial,
                                                              double* __restrict__ tot) {
    constexpr int NG = K * (K + 1) / 2 + K;
    constexpr int NW = PROJ_THREADS / 64;
    __shared__ double ws[NW];
    const int q = blockIdx.x;
    double a = 0.0;
#pragma unroll
    for (int u = 0; u < (PROJ_BLOCKS + PROJ_THREADS - 1) / PROJ_THREADS; ++u) {
        const int blk = threadIdx.x + PROJ_THREADS * u;
        if (blk < PROJ_BLOCKS) a += partial[(size_t)blk * NG + q];
    }
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int u = 0; u < NW; ++u) t += ws[u];
        tot[q] = t;
    }
}

// One wavefront solves for beta from the summed Gram entries: lane r holds
// row r of the scaled Gram matrix, the right-looking Cholesky and the two
// triangular solves take the values of other rows by lane reads, and lane k
// ends with beta_k.
template <int K>
__global__ __launch_bounds__(64) void k_proj_solve(const double* __restrict__ tot, ProjState* ps) {
    static_assert(K <= 64 && PROJ_MAX <= 64, "one row per lane");
    const int lane = threadIdx.x;
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
        g[l] *= s * lane_value(s, l);
        if (l == r) g[l] += 1e-13;  // (keeps the factorisation defined for dependent columns)
    }
    // Cholesky, column by column: lane r's g[k] becomes L[r][k] (r >= k)
    bool ok = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double dkk = lane_value(g[k], k);
        ok = ok && dkk > 0.0;
        const double lkk = sqrt(ok ? dkk : 1.0);
        g[k] = r > k ? g[k] / lkk : (r == k ? lkk : g[k]);
#pragma unroll
        for (int m = k + 1; m < K; ++m) {
            const double lmk = lane_value(g[k], m);
            if (r > k) g[m] -= g[k] * lmk;
        }
    }
    // L y = c (lane k keeps y_k in yl), then L^T x = y (lane k ends with x_k)
    double yl = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double yk = lane_value(c, k) / lane_value(g[k], k);
        if (r > k) c -= g[k] * yk;
        if (r == k) yl = yk;
    }
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        const double xk = lane_value(yl, k) / lane_value(g[k], k);
        if (r == k) yl = xk;
#pragma unroll
        for (int m = 0; m < k; ++m) {
            const double lkm = lane_value(g[m], k);
            if (r == m) yl -= lkm * xk;
        }
    }
    if (lane >= PROJ_MAX) return;
    double bk = 0.0;
    if (lane < K) bk = ok ? (isfinite(yl) ? yl : 0.0) * s : (lane == 0 ? 1.0 : 0.0);
    ps->beta[lane] = bk;  // (no factorisation: the newest pressure alone)
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
    double* tot = partial + (size_t)PROJ_BLOCKS * (K * (K + 1) / 2 + K);
    hipLaunchKernelGGL(k_proj_reduce<K>, dim3(K * (K + 1) / 2 + K), dim3(PROJ_THREADS), 0, s,
                       partial, tot);
    hipLaunchKernelGGL(k_proj_solve<K>, dim3(1), dim3(64), 0, s, tot, ps);
    hipLaunchKernelGGL(k_proj_combine<K>, dim3((n + PROJ_THREADS - 1) / PROJ_THREADS),
                       dim3(PROJ_THREADS), 0, s, n, hp, ld, ps, out);
}

// launch_guess_k<k> for the window k (1 <= k <= K)
template <int K>
void launch_guess_upto(hipStream_t s, int k, int n, const double* hp, const double* hw, size_t ld,
                       ProjState* ps, const double* b, double* partial, double* out) {
    if constexpr (K > 1) {
        if (k < K) return launch_guess_upto<K - 1>(s, k, n, hp, hw, ld, ps, b, partial, out);
    }
    launch_guess_k<K>(s, n, hp, hw, ld, ps, b, partial, out);
}

}  // namespace

size_t proj_partial_size() { return (size_t)(PROJ_BLOCKS + 1) * (PROJ_MAX * (PROJ_MAX + 1) / 2 + PROJ_MAX); }

void launch_proj_guess(hipStream_t s, int k, int n, const double* hp, const double* hw, size_t ld,
                       ProjState* ps, const double* b, double* partial, double* out) {
    launch_guess_upto<PROJ_MAX>(s, k, n, hp, hw, ld, ps, b, partial, out);
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
