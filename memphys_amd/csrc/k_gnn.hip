// Learned initial guess (see gnn.hpp): forward, backward and Adam kernels.
//
// Layouts: node features row-major [nv][F], hidden states row-major [nv][64]
// (one 256-byte row per node, so a wavefront reads a neighbour's state with one
// coalesced load).  Weights are read as BF16 and widened to FP32 in LDS; every
// product accumulates in FP32.  Message passing uses the stencil graph in
// SELL-64x4 form with the FP32 values of A~ = A D^-1.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "gnn.hpp"

namespace mp {

namespace {

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

constexpr int H = GNN_H;
constexpr int F = GNN_F;
constexpr int RG = 4;  // rows per block iteration of the dense kernels (4 x 64 threads)

// sum of squares of (p1 - p2) over the node block -> block partials
__global__ __launch_bounds__(256) void k_sq_partial(int nv, const double* __restrict__ p1,
                                                    const double* __restrict__ p2,
                                                    double* __restrict__ part) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nv; i += gridDim.x * 256) {
        double d = p1[i] - p2[i];
        v += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// s = sqrt(mean) with a floor, into scal[0]
__global__ void k_scale_finish(int nb, int nv, const double* __restrict__ part,
                               double* __restrict__ scal) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    s = sqrt(s / (double)nv);
    scal[0] = (s > 1e-300) ? s : 1.0;
}

__global__ __launch_bounds__(256) void k_features(int nv, const double* __restrict__ p1,
                                                  const double* __restrict__ p2,
                                                  const double* __restrict__ resid,
                                                  const double* __restrict__ dinv,
                                                  const double* __restrict__ xyz, int npad, int dim,
                                                  const double* __restrict__ scal,
                                                  float* __restrict__ X) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nv) return;
    const double inv = 1.0 / scal[0];
    X[i * F + 0] = (float)((p1[i] - p2[i]) * inv);
    X[i * F + 1] = (float)(dinv[i] * resid[i] * inv);
    X[i * F + 2] = (float)xyz[i];
    X[i * F + 3] = (float)xyz[(size_t)npad + i];
    X[i * F + 4] = (dim == 3) ? (float)xyz[2 * (size_t)npad + i] : 0.0f;
}

// Z = in1 W1 (+ in2 W2) + b (not stored when Z is null); Hout = relu(Z), and
// its BF16 copy (round to nearest even) when Hbf is not null.
// in1: [nv][K1], in2: [nv][H].
// A block takes tiles of DR rows: it stages the tile's inputs in LDS and each
// thread forms 4 rows x 4 outputs, every output's fma chain in the order
// bias, in1 columns, in2 columns.
constexpr int DR = 64;  // rows per tile of the dense kernel
template <int K1, bool TWO>
__global__ __launch_bounds__(256) void k_dense(int nv, const float* __restrict__ in1,
                                               const float* __restrict__ in2,
                                               const uint16_t* __restrict__ W1,
                                               const uint16_t* __restrict__ W2,
                                               const uint16_t* __restrict__ bias,
                                               float* __restrict__ Z, float* __restrict__ Hout,
                                               uint16_t* __restrict__ Hbf) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    __shared__ f4 w1[K1 * H / 4];
    __shared__ f4 w2[TWO ? H * H / 4 : 1];
    __shared__ f4 bb[H / 4];
    __shared__ float a1[DR][K1 + 1];
    __shared__ float a2[TWO ? DR : 1][H + 1];
    float* w1s = reinterpret_cast<float*>(w1);
    float* w2s = reinterpret_cast<float*>(w2);
    float* bbs = reinterpret_cast<float*>(bb);
    for (int t = threadIdx.x; t < K1 * H; t += 256) w1s[t] = bf2f(W1[t]);
    if (TWO)
        for (int t = threadIdx.x; t < H * H; t += 256) w2s[t] = bf2f(W2[t]);
    if (threadIdx.x < H) bbs[threadIdx.x] = bf2f(bias[threadIdx.x]);
    const int cg = threadIdx.x & 15;  // outputs 4cg..4cg+3
    const int rg = threadIdx.x >> 4;  // tile rows 4rg..4rg+3
    for (int r0 = blockIdx.x * DR; r0 < nv; r0 += gridDim.x * DR) {
        const int nr = min(DR, nv - r0);
        __syncthreads();  // weights written / previous tile's LDS reads done
        for (int t = threadIdx.x; t < nr * K1; t += 256)
            a1[t / K1][t % K1] = in1[(size_t)r0 * K1 + t];
        if (TWO)
            for (int t = threadIdx.x; t < nr * H; t += 256) a2[t / H][t % H] = in2[(size_t)r0 * H + t];
        __syncthreads();
        f4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = bb[cg];
#pragma unroll 4
        for (int k = 0; k < K1; ++k) {
            const f4 w = w1[k * (H / 4) + cg];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float a = a1[4 * rg + i][k];
                acc[i].x = fmaf(a, w.x, acc[i].x);
                acc[i].y = fmaf(a, w.y, acc[i].y);
                acc[i].z = fmaf(a, w.z, acc[i].z);
                acc[i].w = fmaf(a, w.w, acc[i].w);
            }
        }
        if (TWO) {
#pragma unroll 4
            for (int k = 0; k < H; ++k) {
                const f4 w = w2[k * (H / 4) + cg];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float a = a2[4 * rg + i][k];
                    acc[i].x = fmaf(a, w.x, acc[i].x);
                    acc[i].y = fmaf(a, w.y, acc[i].y);
                    acc[i].z = fmaf(a, w.z, acc[i].z);
                    acc[i].w = fmaf(a, w.w, acc[i].w);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * rg + i;
            if (row >= nr) break;
            const size_t o = (size_t)(r0 + row) * H + 4 * cg;
            if (Z) *reinterpret_cast<f4*>(Z + o) = acc[i];  // kept for the backward pass
            const f4 h = {fmaxf(acc[i].x, 0.0f), fmaxf(acc[i].y, 0.0f), fmaxf(acc[i].z, 0.0f),
                          fmaxf(acc[i].w, 0.0f)};
            *reinterpret_cast<f4*>(Hout + o) = h;
            if (Hbf) {
                typedef uint16_t u4 __attribute__((ext_vector_type(4)));
                *reinterpret_cast<u4*>(Hbf + o) = u4{f2bf(h.x), f2bf(h.y), f2bf(h.z), f2bf(h.w)};
            }
        }
    }
}

// Hidden layer on the matrix cores: Z = in1 W1 + in2 W2 + b, Hout = relu(Z)
// and its BF16 copy (Z and the copy not stored when null); in1, in2: [nv][H], W1, W2: [H][H] (BF16, k-major).
// v_mfma_f32_16x16x4f32 keeps every product and sum in FP32.  A wavefront
// takes tiles of 16 rows; the 128 reduction indices k (in1's 64, then in2's)
// are walked in 32 steps of 4, and lane l holds, for step (j, e), row
// l % 16 of the tile at k = 16 j + 4 (l / 16) + e: one 16-byte load of four
// consecutive inputs per lane and j.  The weights sit in LDS in the same
// fragment order (wf[j][n-block][lane][e]), so each (j, n-block) is one
// 16-byte LDS read.  Accumulators start at the bias.
constexpr int MROWS = 16;  // rows per tile of the matrix-core dense kernel
__global__ __launch_bounds__(256) void k_dense_mfma(int nv, const float* __restrict__ in1,
                                                    const float* __restrict__ in2,
                                                    const uint16_t* __restrict__ W1,
                                                    const uint16_t* __restrict__ W2,
                                                    const uint16_t* __restrict__ bias,
                                                    float* __restrict__ Z,
                                                    float* __restrict__ Hout,
                                                    uint16_t* __restrict__ Hbf) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    static_assert(H == 64, "fragment layout assumes 64 hidden units");
    __shared__ f4 wf[8 * 4 * 64];
    for (int t = threadIdx.x; t < 8 * 4 * 64; t += 256) {
        const int lane = t & 63, nb = (t >> 6) & 3, j = t >> 8;
        const int n = 16 * nb + (lane & 15);
        f4 v;
        for (int e = 0; e < 4; ++e) {
            const int k = 16 * j + 4 * (lane >> 4) + e;
            v[e] = bf2f(k < H ? W1[k * H + n] : W2[(k - H) * H + n]);
        }
        wf[t] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int q = lane >> 4, c = lane & 15;
    float bcol[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) bcol[nb] = bf2f(bias[16 * nb + c]);
    const int ntiles = (nv + MROWS - 1) / MROWS;
    for (int tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * 4) {
        const int r0 = tile * MROWS;
        const int row = r0 + c;
        f4 a[8];
        if (row < nv) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a[j] = *reinterpret_cast<const f4*>(in1 + (size_t)row * H + 16 * j + 4 * q);
                a[j + 4] = *reinterpret_cast<const f4*>(in2 + (size_t)row * H + 16 * j + 4 * q);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        f4 acc[4];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) acc[nb] = f4{bcol[nb], bcol[nb], bcol[nb], bcol[nb]};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const f4 b = wf[(j * 4 + nb) * 64 + lane];
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j].x, b.x, acc[nb], 0, 0, 0);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j].y, b.y, acc[nb], 0, 0, 0);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j].z, b.z, acc[nb], 0, 0, 0);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j].w, b.w, acc[nb], 0, 0, 0);
            }
        }
        // lane holds rows r0 + 4q + i, columns 16 nb + c
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = r0 + 4 * q + i;
            if (r >= nv) break;
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const size_t o = (size_t)r * H + 16 * nb + c;
                if (Z) Z[o] = acc[nb][i];
                const float h = fmaxf(acc[nb][i], 0.0f);
                Hout[o] = h;
                if (Hbf) Hbf[o] = f2bf(h);
            }
        }
    }
}

// out[i][k] = sum_c dZ[i][c] W[k][c] (+ out[i][k] if accumulate); W: [K][H]
template <int K>
__global__ __launch_bounds__(256) void k_dense_t(int nv, const float* __restrict__ dZ,
                                                 const uint16_t* __restrict__ W,
                                                 float* __restrict__ out, int accumulate) {
    __shared__ float w[K * H];
    for (int t = threadIdx.x; t < K * H; t += 256) w[t] = bf2f(W[t]);
    __syncthreads();
    const int k = threadIdx.x & 63;
    if (k >= K) return;
    for (int row = blockIdx.x * RG + (threadIdx.x >> 6); row < nv; row += gridDim.x * RG) {
        const float* d = dZ + (size_t)row * H;
        float acc = accumulate ? out[(size_t)row * K + k] : 0.0f;
#pragma unroll 8
        for (int c = 0; c < H; ++c) acc = fmaf(d[c], w[k * H + c], acc);
        out[(size_t)row * K + k] = acc;
    }
}

// m[i][:] = sum_j a_ij h[j][:] over a SELL-64x4 operator, h in FP32 or as
// BF16 numbers (TH = uint16_t; the products and sums stay FP32).  A wavefront
// takes four consecutive rows, sixteen lanes per row, each lane four features
// (one 16- or 8-byte gather per neighbour), so one gather instruction fetches
// four neighbour rows.  accumulate: m += instead of m =.  A row's slots come in
// groups of four (one 16-byte load of columns and one of values, the row's
// own lanes of the slice); the gathers of two groups are issued before their
// fmas, which keep the order of the row's entries.
template <typename TH>
__global__ __launch_bounds__(256) void k_aggregate(SellView sv, const float* __restrict__ val,
                                                   const TH* __restrict__ h,
                                                   float* __restrict__ m, int accumulate) {
    typedef int i4 __attribute__((ext_vector_type(4)));
    typedef float f4 __attribute__((ext_vector_type(4)));
    static_assert(H == 64, "sixteen lanes of four features per row");
    const int row = (blockIdx.x * 4 + (int)(threadIdx.x >> 6)) * 4 + (int)((threadIdx.x >> 4) & 3);
    if (row >= sv.nv) return;
    const int fq = 4 * (threadIdx.x & 15);  // features fq..fq+3
    const int base = sell_base(sv.sptr, row);
    const int len = sv.len[row];
    f4 acc = accumulate ? *reinterpret_cast<const f4*>(m + (size_t)row * H + fq)
                        : f4{0.0f, 0.0f, 0.0f, 0.0f};
    auto gather = [&](int c) {
        if constexpr (sizeof(TH) == 2) {
            typedef uint16_t u4 __attribute__((ext_vector_type(4)));
            const u4 b = *reinterpret_cast<const u4*>(h + (size_t)c * H + fq);
            return f4{bf2f(b.x), bf2f(b.y), bf2f(b.z), bf2f(b.w)};
        } else {
            return *reinterpret_cast<const f4*>(h + (size_t)c * H + fq);
        }
    };
    auto fma4 = [](float a, const f4& x, f4 y) {
        y.x = fmaf(a, x.x, y.x);
        y.y = fmaf(a, x.y, y.y);
        y.z = fmaf(a, x.z, y.z);
        y.w = fmaf(a, x.w, y.w);
        return y;
    };
    int j0 = 0;
    for (; j0 + 8 <= len; j0 += 8) {
        const int s4 = base + (j0 << 6);
        const i4 c0 = *reinterpret_cast<const i4*>(sv.col + s4);
        const i4 c1 = *reinterpret_cast<const i4*>(sv.col + s4 + 256);
        const f4 v0 = *reinterpret_cast<const f4*>(val + s4);
        const f4 v1 = *reinterpret_cast<const f4*>(val + s4 + 256);
        const f4 x0 = gather(c0.x), x1 = gather(c0.y), x2 = gather(c0.z), x3 = gather(c0.w);
        const f4 x4 = gather(c1.x), x5 = gather(c1.y), x6 = gather(c1.z), x7 = gather(c1.w);
        acc = fma4(v0.x, x0, acc);
        acc = fma4(v0.y, x1, acc);
        acc = fma4(v0.z, x2, acc);
        acc = fma4(v0.w, x3, acc);
        acc = fma4(v1.x, x4, acc);
        acc = fma4(v1.y, x5, acc);
        acc = fma4(v1.z, x6, acc);
        acc = fma4(v1.w, x7, acc);
    }
    for (; j0 < len; j0 += 4) {
        const int s4 = base + (j0 << 6);
        const i4 cc = *reinterpret_cast<const i4*>(sv.col + s4);
        const f4 vv = *reinterpret_cast<const f4*>(val + s4);
        const int n = min(4, len - j0);  // entries of this group inside the row
        acc = fma4(vv.x, gather(cc.x), acc);
        if (n > 1) acc = fma4(vv.y, gather(cc.y), acc);
        if (n > 2) acc = fma4(vv.z, gather(cc.z), acc);
        if (n > 3) acc = fma4(vv.w, gather(cc.w), acc);
    }
    *reinterpret_cast<f4*>(m + (size_t)row * H + fq) = acc;
}

// d_i = h_i . wo + bo; one wavefront per row
__global__ __launch_bounds__(256) void k_output(int nv, const float* __restrict__ h,
                                                const uint16_t* __restrict__ wo,
                                                const uint16_t* __restrict__ bo,
                                                float* __restrict__ d) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nv) return;
    const int c = threadIdx.x & 63;
    float v = h[(size_t)row * H + c] * bf2f(wo[c]);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (c == 0) d[row] = v + bf2f(bo[0]);
}

// guess = base (the extrapolated pressure) + s d on the node rows
__global__ __launch_bounds__(256) void k_guess(int nv, const double* __restrict__ base,
                                               const float* __restrict__ d,
                                               const double* __restrict__ scal,
                                               double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nv) return;
    out[i] = base[i] + scal[0] * (double)d[i];
}

// e_i = d_i - t_i with t = (p* - base)/s; dd_i = 2 e_i / nv; partial sums of e^2
__global__ __launch_bounds__(256) void k_loss(int nv, const float* __restrict__ d,
                                              const double* __restrict__ base,
                                              const double* __restrict__ pt,
                                              const double* __restrict__ scal,
                                              float* __restrict__ dd, double* __restrict__ part) {
    __shared__ double sh[4];
    double acc = 0.0;
    const double inv = 1.0 / scal[0];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nv; i += gridDim.x * 256) {
        const double t = (pt[i] - base[i]) * inv;
        const double e = (double)d[i] - t;
        acc += e * e;
        dd[i] = (float)(2.0 * e / (double)nv);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_loss_finish(int nb, int nv, const double* __restrict__ part,
                              double* __restrict__ scal) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    scal[1] = s / (double)nv;
}

// dH[i][c] = dd_i wo[c]
__global__ __launch_bounds__(256) void k_outer_wo(int nv, const float* __restrict__ dd,
                                                  const uint16_t* __restrict__ wo,
                                                  float* __restrict__ dH) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)nv * H) return;
    dH[t] = dd[t / H] * bf2f(wo[t % H]);
}

// dZ = dH * (Z > 0)
__global__ __launch_bounds__(256) void k_relu_back(size_t n, const float* __restrict__ dH,
                                                   const float* __restrict__ Z,
                                                   float* __restrict__ dZ) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) dZ[t] = (Z[t] > 0.0f) ? dH[t] : 0.0f;
}

// Weight gradients: part[b][k][c] = sum over this block's rows of A[i][k] B[i][c]
// for k < KA, plus row KA = sum B[i][c] (bias) when BIAS.  256 threads own the
// (KA+BIAS) x KB outputs; rows are staged through LDS 16 at a time.
template <int KA, int KB, bool BIAS>
__global__ __launch_bounds__(256) void k_wgrad_partial(int nv, const float* __restrict__ A,
                                                       const float* __restrict__ B,
                                                       float* __restrict__ part) {
    constexpr int KR = KA + (BIAS ? 1 : 0);
    constexpr int NOUT = KR * KB;
    constexpr int PER = (NOUT + 255) / 256;
    constexpr int TR = 16;
    __shared__ float sa[TR][KA > 0 ? KA : 1];
    __shared__ float sb[TR][KB];
    float acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = 0.0f;
    const int rows_per_block = (nv + gridDim.x - 1) / gridDim.x;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(nv, r0 + rows_per_block);
    for (int t0 = r0; t0 < r1; t0 += TR) {
        const int nr = min(TR, r1 - t0);
        for (int t = threadIdx.x; t < TR * KA; t += 256) {
            const int rr = t / KA, kk = t - rr * KA;
            sa[rr][kk] = (rr < nr) ? A[(size_t)(t0 + rr) * KA + kk] : 0.0f;
        }
        for (int t = threadIdx.x; t < TR * KB; t += 256) {
            const int rr = t / KB, kk = t - rr * KB;
            sb[rr][kk] = (rr < nr) ? B[(size_t)(t0 + rr) * KB + kk] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int o = threadIdx.x + q * 256;
            if (o < NOUT) {
                const int k = o / KB, c = o - k * KB;
                float a = acc[q];
                for (int rr = 0; rr < TR; ++rr) {
                    const float av = (k < KA) ? sa[rr][k] : 1.0f;
                    a = fmaf(av, sb[rr][c], a);
                }
                acc[q] = a;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int o = threadIdx.x + q * 256;
        if (o < NOUT) part[(size_t)blockIdx.x * NOUT + o] = acc[q];
    }
}

// grad[off + o] = sum_b part[b][o] (block order); bias row goes to bias_off
__global__ __launch_bounds__(256) void k_wgrad_finish(int nb, int nout, int kb, int ka,
                                                      const float* __restrict__ part,
                                                      float* __restrict__ grad, int w_off,
                                                      int bias_off) {
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= nout) return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * nout + o];
    const int k = o / kb, c = o - k * kb;
    if (k < ka)
        grad[w_off + o] = s;
    else
        grad[bias_off + c] = s;
}

__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ th,
                                              const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, uint16_t* __restrict__ thbf,
                                              float lr, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float t = th[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    th[i] = t;
    thbf[i] = f2bf(t);
}

__global__ __launch_bounds__(256) void k_to_bf(int n, const float* __restrict__ th,
                                               uint16_t* __restrict__ bf) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) bf[i] = f2bf(th[i]);
}

inline int blocks(size_t n) { return (int)((n + 255) / 256); }
constexpr int DENSE_GRID = 2048;
constexpr int WG_GRID = 256;

}  // namespace

void GuessNet::setup(hipStream_t st, const SellView& view, const float* a32_slots,
                     int nslots_total) {
    sv = view;
    nv = view.nv;
    a32 = a32_slots;
    // transposed node-block operator, built on the host once (needed by backward)
    const int nsl = view.nslices;
    std::vector<int> hsptr(nsl + 1), hlen((size_t)nsl * SLICE), hcol(nslots_total);
    std::vector<float> hval(nslots_total);
    MP_HIP(hipMemcpy(hsptr.data(), view.sptr, sizeof(int) * (nsl + 1), hipMemcpyDeviceToHost));
    MP_HIP(hipMemcpy(hlen.data(), view.len, sizeof(int) * hlen.size(), hipMemcpyDeviceToHost));
    MP_HIP(hipMemcpy(hcol.data(), view.col, sizeof(int) * nslots_total, hipMemcpyDeviceToHost));
    MP_HIP(hipMemcpy(hval.data(), a32_slots, sizeof(float) * nslots_total, hipMemcpyDeviceToHost));
    std::vector<int> tcnt(nv + 1, 0);
    for (int r = 0; r < nv; ++r)
        for (int j = 0; j < hlen[r]; ++j) tcnt[hcol[sell_base(hsptr.data(), r) + sell_off(j)] + 1]++;
    std::vector<int> trp(nv + 1, 0);
    for (int i = 0; i < nv; ++i) trp[i + 1] = trp[i] + tcnt[i + 1];
    std::vector<int> tc(trp[nv]);
    std::vector<float> tv(trp[nv]);
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (int r = 0; r < nv; ++r)
        for (int j = 0; j < hlen[r]; ++j) {
            const int slot = sell_base(hsptr.data(), r) + sell_off(j);
            const int c = hcol[slot];
            tc[fill[c]] = r;
            tv[fill[c]] = hval[slot];
            fill[c]++;
        }
    std::vector<int> tsp(nsl + 1, 0), tl((size_t)nsl * SLICE, 0);
    for (int i = 0; i < nv; ++i) tl[i] = trp[i + 1] - trp[i];
    for (int s = 0; s < nsl; ++s) {
        int w = 0;
        for (int l = 0; l < SLICE; ++l) w = std::max(w, tl[s * SLICE + l]);
        tsp[s + 1] = tsp[s] + sell_width4(w) * SLICE;
    }
    std::vector<int> tcol_h(std::max(1, tsp[nsl]), 0);
    std::vector<float> tval_h(std::max(1, tsp[nsl]), 0.0f);
    for (int i = 0; i < nv; ++i)
        for (int j = 0; j < tl[i]; ++j) {
            const int slot = sell_base(tsp.data(), i) + sell_off(j);
            tcol_h[slot] = tc[trp[i] + j];
            tval_h[slot] = tv[trp[i] + j];
        }
    tsptr.alloc(tsp.size());
    tlen.alloc(tl.size());
    tcol.alloc(tcol_h.size());
    tval.alloc(tval_h.size());
    MP_HIP(hipMemcpy(tsptr.p, tsp.data(), sizeof(int) * tsp.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(tlen.p, tl.data(), sizeof(int) * tl.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(tcol.p, tcol_h.data(), sizeof(int) * tcol_h.size(), hipMemcpyHostToDevice));
    MP_HIP(hipMemcpy(tval.p, tval_h.data(), sizeof(float) * tval_h.size(), hipMemcpyHostToDevice));
    tsv.nv = nv;
    tsv.nslices = nsl;
    tsv.sptr = tsptr.p;
    tsv.col = tcol.p;
    tsv.len = tlen.p;
    const size_t nh = (size_t)nv * H;
    X.alloc((size_t)nv * F);
    for (int l = 0; l <= GNN_L; ++l) {
        Z[l].alloc(nh);
        Hh[l].alloc(nh);
        Mm[l].alloc(nh);
        if (l < GNN_L) Hb[l].alloc(nh);
    }
    out.alloc(nv);
    dbuf.alloc(nv);
    dH.alloc(nh);
    dZ.alloc(nh);
    dTmp.alloc(nh);
    theta.alloc(GNN_NPARAM);
    grad.alloc(GNN_NPARAM);
    adam_m.alloc(GNN_NPARAM);
    adam_v.alloc(GNN_NPARAM);
    theta_bf.alloc(GNN_NPARAM);
    scal.alloc(8);
    partial.alloc(4096);
    wpart_blocks = WG_GRID;
    wpart.alloc((size_t)WG_GRID * (H + 1) * H);
    MP_HIP(hipMemsetAsync(adam_m.p, 0, sizeof(float) * GNN_NPARAM, st));
    MP_HIP(hipMemsetAsync(adam_v.p, 0, sizeof(float) * GNN_NPARAM, st));
    MP_HIP(hipMemsetAsync(grad.p, 0, sizeof(float) * GNN_NPARAM, st));
    adam_t = 0;
}

void GuessNet::init_params(hipStream_t st, uint64_t seed) {
    std::vector<float> h(GNN_NPARAM, 0.0f);
    std::mt19937_64 rng(seed);
    auto uni = [&](float a) {
        return (float)(((double)(rng() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * a);
    };
    const float a0 = std::sqrt(6.0f / (GNN_F + GNN_H));
    for (int t = 0; t < GNN_F * GNN_H; ++t) h[GNN_W0 + t] = uni(a0);
    const float a1 = std::sqrt(6.0f / (2 * GNN_H)) * 0.5f;
    for (int l = 0; l < GNN_L; ++l) {
        const int o = GNN_LAYER + l * GNN_LAYER_SIZE;
        for (int t = 0; t < 2 * GNN_H * GNN_H; ++t) h[o + t] = uni(a1);
    }
    // output layer starts at zero: the untrained guess is the extrapolation
    set_params(st, h.data());
}

void GuessNet::set_params(hipStream_t st, const float* hp) {
    MP_HIP(hipMemcpyAsync(theta.p, hp, sizeof(float) * GNN_NPARAM, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_to_bf, dim3(blocks(GNN_NPARAM)), dim3(256), 0, st, GNN_NPARAM, theta.p,
                       theta_bf.p);
    MP_HIP(hipGetLastError());
    MP_HIP(hipMemsetAsync(adam_m.p, 0, sizeof(float) * GNN_NPARAM, st));
    MP_HIP(hipMemsetAsync(adam_v.p, 0, sizeof(float) * GNN_NPARAM, st));
    adam_t = 0;
    MP_HIP(hipStreamSynchronize(st));
}

void GuessNet::get_params(hipStream_t st, float* hp) {
    MP_HIP(hipMemcpyAsync(hp, theta.p, sizeof(float) * GNN_NPARAM, hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
}

void GuessNet::features(hipStream_t st, const double* p1, const double* p2, const double* resid,
                        const double* dinv, const double* xyz_soa, int npad, int dim) {
    const int nb = 256;
    hipLaunchKernelGGL(k_sq_partial, dim3(nb), dim3(256), 0, st, nv, p1, p2, partial.p);
    hipLaunchKernelGGL(k_scale_finish, dim3(1), dim3(64), 0, st, nb, nv, partial.p, scal.p);
    hipLaunchKernelGGL(k_features, dim3(blocks(nv)), dim3(256), 0, st, nv, p1, p2, resid, dinv,
                       xyz_soa, npad, dim, scal.p, X.p);
    MP_HIP(hipGetLastError());
}

void GuessNet::forward(hipStream_t st, bool keep_pre) {
    const uint16_t* P = theta_bf.p;
    const int g = std::min(DENSE_GRID, (nv + DR - 1) / DR);
    const int gm = std::min(DENSE_GRID, (nv + 4 * MROWS - 1) / (4 * MROWS));
    // each hidden state the next layer aggregates also gets a BF16 copy, which
    // the aggregation gathers (half the bytes of the FP32 rows)
    hipLaunchKernelGGL((k_dense<F, false>), dim3(g), dim3(256), 0, st, nv, X.p, nullptr,
                       P + GNN_W0, nullptr, P + GNN_B0, keep_pre ? Z[0].p : nullptr, Hh[0].p,
                       Hb[0].p);
    for (int l = 1; l <= GNN_L; ++l) {
        const int o = GNN_LAYER + (l - 1) * GNN_LAYER_SIZE;
        hipLaunchKernelGGL(k_aggregate<uint16_t>, dim3((nv + 15) / 16), dim3(256), 0, st, sv, a32,
                           (const uint16_t*)Hb[l - 1].p, Mm[l].p, 0);
        hipLaunchKernelGGL(k_dense_mfma, dim3(gm), dim3(256), 0, st, nv, Hh[l - 1].p, Mm[l].p,
                           P + o, P + o + H * H, P + o + 2 * H * H, keep_pre ? Z[l].p : nullptr,
                           Hh[l].p, l < GNN_L ? Hb[l].p : nullptr);
    }
    hipLaunchKernelGGL(k_output, dim3((nv + 3) / 4), dim3(256), 0, st, nv, Hh[GNN_L].p, P + GNN_WO,
                       P + GNN_BO, out.p);
    MP_HIP(hipGetLastError());
}

void GuessNet::guess(hipStream_t st, const double* base, double* dst) {
    hipLaunchKernelGGL(k_guess, dim3(blocks(nv)), dim3(256), 0, st, nv, base, out.p, scal.p, dst);
    MP_HIP(hipGetLastError());
}

void GuessNet::backward(hipStream_t st, const double* base, const double* pt) {
    const uint16_t* P = theta_bf.p;
    const int nb = 256;
    const size_t nh = (size_t)nv * H;
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, st, nv, out.p, base, pt, scal.p, dbuf.p,
                       partial.p);
    hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(64), 0, st, nb, nv, partial.p, scal.p);
    // output layer: grad wo, bo; dH3
    hipLaunchKernelGGL((k_wgrad_partial<H, 1, true>), dim3(WG_GRID), dim3(256), 0, st, nv,
                       Hh[GNN_L].p, dbuf.p, wpart.p);
    hipLaunchKernelGGL(k_wgrad_finish, dim3(blocks(H + 1)), dim3(256), 0, st, WG_GRID, H + 1, 1, H,
                       wpart.p, grad.p, GNN_WO, GNN_BO);
    hipLaunchKernelGGL(k_outer_wo, dim3(blocks(nh)), dim3(256), 0, st, nv, dbuf.p, P + GNN_WO, dH.p);
    const int g = std::min(DENSE_GRID, (nv + RG - 1) / RG);
    for (int l = GNN_L; l >= 1; --l) {
        const int o = GNN_LAYER + (l - 1) * GNN_LAYER_SIZE;
        hipLaunchKernelGGL(k_relu_back, dim3(blocks(nh)), dim3(256), 0, st, nh, dH.p, Z[l].p, dZ.p);
        // dWs = h_{l-1}^T dZ (+ db), dWn = m_l^T dZ
        hipLaunchKernelGGL((k_wgrad_partial<H, H, true>), dim3(WG_GRID), dim3(256), 0, st, nv,
                           Hh[l - 1].p, dZ.p, wpart.p);
        hipLaunchKernelGGL(k_wgrad_finish, dim3(blocks((H + 1) * H)), dim3(256), 0, st, WG_GRID,
                           (H + 1) * H, H, H, wpart.p, grad.p, o, o + 2 * H * H);
        hipLaunchKernelGGL((k_wgrad_partial<H, H, false>), dim3(WG_GRID), dim3(256), 0, st, nv,
                           Mm[l].p, dZ.p, wpart.p);
        hipLaunchKernelGGL(k_wgrad_finish, dim3(blocks(H * H)), dim3(256), 0, st, WG_GRID, H * H,
                           H, H, wpart.p, grad.p, o + H * H, 0);
        // dH_{l-1} = dZ Ws^T + A~^T (dZ Wn^T)
        hipLaunchKernelGGL((k_dense_t<H>), dim3(g), dim3(256), 0, st, nv, dZ.p, P + o + H * H,
                           dTmp.p, 0);
        hipLaunchKernelGGL((k_dense_t<H>), dim3(g), dim3(256), 0, st, nv, dZ.p, P + o, dH.p, 0);
        hipLaunchKernelGGL(k_aggregate<float>, dim3((nv + 15) / 16), dim3(256), 0, st, tsv, tval.p,
                           (const float*)dTmp.p, dH.p, 1);
    }
    hipLaunchKernelGGL(k_relu_back, dim3(blocks(nh)), dim3(256), 0, st, nh, dH.p, Z[0].p, dZ.p);
    hipLaunchKernelGGL((k_wgrad_partial<F, H, true>), dim3(WG_GRID), dim3(256), 0, st, nv, X.p,
                       dZ.p, wpart.p);
    hipLaunchKernelGGL(k_wgrad_finish, dim3(blocks((F + 1) * H)), dim3(256), 0, st, WG_GRID,
                       (F + 1) * H, H, F, wpart.p, grad.p, GNN_W0, GNN_B0);
    MP_HIP(hipGetLastError());
}

void GuessNet::adam_step(hipStream_t st) {
    adam_t += 1;
    const float c1 = 1.0f - std::pow(0.9f, (float)adam_t);
    const float c2 = 1.0f - std::pow(0.999f, (float)adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(GNN_NPARAM)), dim3(256), 0, st, GNN_NPARAM, theta.p,
                       grad.p, adam_m.p, adam_v.p, theta_bf.p, lr, c1, c2);
    MP_HIP(hipGetLastError());
}

double GuessNet::loss_host(hipStream_t st) {
    double h[2];
    MP_HIP(hipMemcpyAsync(h, scal.p, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
    return h[1];
}

}  // namespace mp
