// Smoothed-aggregation V(1,1)-cycle on gfx950: FP32 arithmetic and vectors,
// BF16 matrix values (a preconditioner needs no more: the V-cycle with BF16
// matrices takes as many BiCGSTAB iterations as with FP32 ones), which halves
// the value traffic of every product in the cycle.
//
// Per level l with right-hand side b_l (zero initial guess):
//   pre-smooth + residual  x_l = w D^-1 b_l,  res_l = b_l - A_l x_l    (one kernel)
//   restrict               b_{l+1} = R_l res_l
//   ... coarser levels, dense solve on the coarsest ...
//   prolong                x_l += P_l y_{l+1}
//   post-smooth            y_l = x_l + w D^-1 (b_l - A_l x_l)
// Level 0 reads its right-hand side from the Krylov vector (FP64) and writes
// the FP64 result; the border unknown of the full-Neumann system passes
// through unchanged.
//
// With three or more levels the cycle below level 1 runs from matrices that
// compose its steps, built at set-up (the same linear map, one rounding of
// the composed values instead of one per factor), so each coarse level costs
// one product on the way down and one on the way up:
//   restrict + pre-smooth  b_l = R_{l-1} res_{l-1},  res_l = Q_l res_{l-1},
//                          Q_l = R_{l-1} - A_l D_l R_{l-1}        (levels 2 .. L-2)
//   coarsest solve         y_{L-1} = (G R_{L-2}) res_{L-2}        (dense rows; one
//                          problem: a batch's many small problems keep
//                          R_{L-2} and the block-diagonal G, fewer bytes)
//   prolong + post-smooth  y_l = D_l (b_l + res_l) + T_l y_{l+1},
//                          T_l = P_l - D_l A_l P_l                (levels 1 .. L-2)
// (with x_l = D_l b_l after pre-smoothing: x_l + P_l y + D_l (b_l - A_l (x_l +
// P_l y)) = D_l b_l + D_l res_l + (P_l - D_l A_l P_l) y).  Levels 0 and 1 keep
// the separate restriction and pre-smoothing (Q_1 would be about 25 times
// denser than A_1), and level 0 the separate prolongation and post-smoothing.  A matrix is SELL-64x4 (one wavefront per slice of 64
// rows, four neighbours of a row side by side, padding slots with value 0)
// or CSR with several lanes per row, whichever suits its row lengths.
#include "amg.hpp"
#include "krylov.hpp"

#include <chrono>
#include <cstring>
#include <type_traits>

namespace mp {

namespace {

enum { OP_PRE = 0, OP_SPMV = 1, OP_POST = 2, OP_PRES = 3, OP_RQ = 4, OP_TPOST = 5 };

__host__ __device__ inline int nblk_rows(int nslices) { return (nslices * SLICE + BLOCK - 1) / BLOCK; }

__device__ __forceinline__ bool stopped(const int* done) { return done && *done; }

struct MatView {
    int n, nslices;
    const int* ptr;
    const int* col;
    const uint16_t* val;    // BF16
    const uint16_t* col16;  // compressed columns (SELL-64x4 only; may be null)
    const int* cbase;
    const uint16_t* val2;   // second BF16 value array in the same slots (OP_RQ)
};

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

// FP32 -> BF16, round to nearest even (finite values)
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof(u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(BLOCK) void k_to_bf16(long n, const float* __restrict__ in,
                                                   uint16_t* __restrict__ out) {
    const long i = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// Entry of the vector a row gathers: D^-1 b for the pre-smoothing product,
// the input vector otherwise.
template <int OP, typename TB>
__device__ __forceinline__ float gathered(int c, const TB* __restrict__ b, const float* xin,
                                          const float* __restrict__ d) {
    if constexpr (OP == OP_PRE)
        return d[c] * (float)b[c];
    else
        return xin[c];
}

// One multigrid product with its epilogue, in either layout (G == 0: SELL-64x4,
// else CSR with G lanes per row):
//   OP_PRE : xout = D^-1 b,  res = b - A (D^-1 b)
//   OP_PRES: res = b - A xin, with xin = D^-1 b already formed
//   OP_SPMV: out = base + M xin (base may be null and may alias out)
//   OP_POST: out = xin + D^-1 (b - A xin)
//   OP_RQ  : xout = M2 xin, res = M xin (M = Q_l in val, M2 = R_{l-1} in val2;
//            CSR layouts only)
//   OP_TPOST: out = D^-1 (b + base) + M xin (M = T_l, base = res_l)
// Level 0 also copies the border unknown (border_in -> border_out).
template <int G, int OP, typename TB, typename TO>
__global__ __launch_bounds__(BLOCK) void k_amg_op(MatView m, const float* __restrict__ dinv,
                                                  const TB* __restrict__ b, const float* xin,
                                                  const float* base, float* __restrict__ xout,
                                                  float* __restrict__ res, TO* out,
                                                  const double* __restrict__ border_in,
                                                  double* __restrict__ border_out,
                                                  const int* done, float* __restrict__ out32) {
    if (stopped(done)) return;
    float acc = 0.0f, acc2 = 0.0f;
    int i;
    if constexpr (G == 0) {
        static_assert(OP != OP_RQ, "two value arrays: CSR layouts only");
        const int lb = xcd_block(nblk_rows(m.nslices));
        if (lb < 0) return;
        if (border_out && lb == 0 && threadIdx.x == 0) *border_out = *border_in;
        i = lb * BLOCK + threadIdx.x;
        const int s = __builtin_amdgcn_readfirstlane(i >> 6);
        if (s >= m.nslices) return;
        acc = sell4_accumulate<4, float>(
            m.ptr, m.col, m.col16, m.cbase, m.val, s, threadIdx.x & 63, [](uint16_t v) { return bf2f(v); },
            [=](int c) { return gathered<OP>(c, b, xin, dinv); });
        if (i >= m.n) return;
    } else {
        constexpr int RPB = BLOCK / G;
        const int lb = xcd_block((m.n + RPB - 1) / RPB);
        if (lb < 0) return;
        if (border_out && lb == 0 && threadIdx.x == 0) *border_out = *border_in;
        i = lb * RPB + (int)(threadIdx.x / G);
        const int sub = threadIdx.x % G;
        if (i < m.n) {
            // each lane sums entries sub, sub + G, ... of the row in order; the
            // columns and values of four of them are loaded before their gathers
            const int t1 = m.ptr[i + 1];
            int t = m.ptr[i] + sub;
            for (; t + 3 * G < t1; t += 4 * G) {
                const int c0 = m.col[t], c1 = m.col[t + G], c2 = m.col[t + 2 * G],
                          c3 = m.col[t + 3 * G];
                const uint16_t v0 = m.val[t], v1 = m.val[t + G], v2 = m.val[t + 2 * G],
                               v3 = m.val[t + 3 * G];
                const float x0 = gathered<OP>(c0, b, xin, dinv), x1 = gathered<OP>(c1, b, xin, dinv),
                            x2 = gathered<OP>(c2, b, xin, dinv), x3 = gathered<OP>(c3, b, xin, dinv);
                acc = fmaf(bf2f(v0), x0, acc);
                acc = fmaf(bf2f(v1), x1, acc);
                acc = fmaf(bf2f(v2), x2, acc);
                acc = fmaf(bf2f(v3), x3, acc);
                if constexpr (OP == OP_RQ) {
                    const uint16_t u0 = m.val2[t], u1 = m.val2[t + G], u2 = m.val2[t + 2 * G],
                                   u3 = m.val2[t + 3 * G];
                    acc2 = fmaf(bf2f(u0), x0, acc2);
                    acc2 = fmaf(bf2f(u1), x1, acc2);
                    acc2 = fmaf(bf2f(u2), x2, acc2);
                    acc2 = fmaf(bf2f(u3), x3, acc2);
                }
            }
            if (t < t1) {
                // the lane's last one to three entries, loads issued together
                // (an absent entry re-reads entry t and its term is skipped)
                const bool h1 = t + G < t1, h2 = t + 2 * G < t1;
                const int ta = h1 ? t + G : t, tb = h2 ? t + 2 * G : t;
                const int c0 = m.col[t], c1 = m.col[ta], c2 = m.col[tb];
                const uint16_t v0 = m.val[t], v1 = m.val[ta], v2 = m.val[tb];
                const float x0 = gathered<OP>(c0, b, xin, dinv), x1 = gathered<OP>(c1, b, xin, dinv),
                            x2 = gathered<OP>(c2, b, xin, dinv);
                acc = fmaf(bf2f(v0), x0, acc);
                if (h1) acc = fmaf(bf2f(v1), x1, acc);
                if (h2) acc = fmaf(bf2f(v2), x2, acc);
                if constexpr (OP == OP_RQ) {
                    const uint16_t u0 = m.val2[t], u1 = m.val2[ta], u2 = m.val2[tb];
                    acc2 = fmaf(bf2f(u0), x0, acc2);
                    if (h1) acc2 = fmaf(bf2f(u1), x1, acc2);
                    if (h2) acc2 = fmaf(bf2f(u2), x2, acc2);
                }
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            acc += __shfl_xor(acc, off);
            if constexpr (OP == OP_RQ) acc2 += __shfl_xor(acc2, off);
        }
        if (sub != 0 || i >= m.n) return;
    }
    if constexpr (OP == OP_PRE) {
        const float bi = (float)b[i];
        xout[i] = dinv[i] * bi;
        res[i] = bi - acc;
    } else if constexpr (OP == OP_PRES) {
        res[i] = (float)b[i] - acc;
    } else if constexpr (OP == OP_SPMV) {
        out[i] = (TO)(base ? base[i] + acc : acc);
    } else if constexpr (OP == OP_RQ) {
        xout[i] = acc2;
        res[i] = acc;
    } else if constexpr (OP == OP_TPOST) {
        out[i] = (TO)(dinv[i] * ((float)b[i] + base[i]) + acc);
    } else {
        const float y = xin[i] + dinv[i] * ((float)b[i] - acc);
        out[i] = (TO)y;
        if (out32) out32[i] = y;  // the same values in FP32 (gathered by the Krylov products)
    }
}

// y = G b on the coarsest level: one wavefront per row.
template <typename TB, typename TO>
__global__ __launch_bounds__(BLOCK) void k_amg_dense(int nc, const float* __restrict__ G,
                                                     const TB* __restrict__ b,
                                                     TO* __restrict__ y,
                                                     const double* __restrict__ border_in,
                                                     double* __restrict__ border_out,
                                                     const int* done, float* __restrict__ y32) {
    if (stopped(done)) return;
    if (border_out && blockIdx.x == 0 && threadIdx.x == 0) *border_out = *border_in;
    const int row = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)));
    if (row >= nc) return;
    const int lane = threadIdx.x & 63;
    float acc = 0.0f;
    for (int j = lane; j < nc; j += 64) acc = fmaf(G[(size_t)row * nc + j], (float)b[j], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        y[row] = (TO)acc;
        if (y32) y32[row] = acc;
    }
}

// y = blockdiag(G_q) b on the coarsest level of a batch: one wavefront per
// row; row r's block starts at row first[r], has size[r] rows, and its row of
// the block's dense inverse starts at G + data[r].
template <typename TB, typename TO>
__global__ __launch_bounds__(BLOCK) void k_amg_blockdense(int nc, const float* __restrict__ G,
                                                          const int* __restrict__ first,
                                                          const int* __restrict__ size,
                                                          const long* __restrict__ data,
                                                          const TB* __restrict__ b,
                                                          TO* __restrict__ y,
                                                          float* __restrict__ y32) {
    const int row = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)));
    if (row >= nc) return;
    const int lane = threadIdx.x & 63;
    const int f = first[row], m = size[row];
    const float* g = G + data[row];
    float acc = 0.0f;
    for (int j = lane; j < m; j += 64) acc = fmaf(g[j], (float)b[f + j], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        y[row] = (TO)acc;
        if (y32) y32[row] = acc;
    }
}

// y[r] = sum_j C[r, j] x[first[r] + j] over j < size[r]: the coarsest solve
// composed with the restriction to it (C = G R_{L-2} per problem, BF16, each
// row starting at data[r], a multiple of 4).  One block per row; each thread
// sums entries 4t .. 4t+3, 4t+4096 .., then a fixed reduction tree.
constexpr int RD_THREADS = 1024;
__global__ __launch_bounds__(RD_THREADS) void k_amg_rowdense(int nrows,
                                                              const uint16_t* __restrict__ C,
                                                              const int* __restrict__ first,
                                                              const int* __restrict__ size,
                                                              const long* __restrict__ data,
                                                              const float* __restrict__ x,
                                                              float* __restrict__ y,
                                                              const int* done) {
    if (stopped(done)) return;
    const int r = blockIdx.x;
    if (r >= nrows) return;
    const int f = first[r], m = size[r];
    const uint16_t* c = C + data[r];
    float acc = 0.0f;
    for (int j = 4 * (int)threadIdx.x; j < m; j += 4 * RD_THREADS) {
        const uint2 v = *reinterpret_cast<const uint2*>(c + j);  // four BF16, 8-byte aligned
        acc = fmaf(__uint_as_float(v.x << 16), x[f + j], acc);
        if (j + 1 < m) acc = fmaf(__uint_as_float(v.x & 0xFFFF0000u), x[f + j + 1], acc);
        if (j + 2 < m) acc = fmaf(__uint_as_float(v.y << 16), x[f + j + 2], acc);
        if (j + 3 < m) acc = fmaf(__uint_as_float(v.y & 0xFFFF0000u), x[f + j + 3], acc);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ float part[RD_THREADS / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = part[0];
        for (int w = 1; w < RD_THREADS / 64; ++w) t += part[w];
        y[r] = t;
    }
}

// out[border row] = in[border row] for every bordered problem of a batch
__global__ void k_border_copy(int n, const int* __restrict__ rows, const double* __restrict__ in,
                              double* __restrict__ out) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t < n) out[rows[t]] = in[rows[t]];
}

template <int G, int OP, typename TB, typename TO>
void launch_g(hipStream_t st, const AmgMat& M, const float* dinv, const TB* b, const float* xin,
              const float* base, float* xout, float* res, TO* out, const double* bin,
              double* bout, const int* done, float* out32) {
    const MatView v{M.n, M.nslices, M.ptr, M.col, M.val, M.col16, M.cbase, M.val2};
    int nb;
    if constexpr (G == 0)
        nb = nblk_rows(M.nslices);
    else
        nb = (M.n + BLOCK / G - 1) / (BLOCK / G);
    hipLaunchKernelGGL((k_amg_op<G, OP, TB, TO>), dim3(grid_xcd(nb)), dim3(BLOCK), 0, st, v, dinv,
                       b, xin, base, xout, res, out, bin, bout, done, out32);
}

template <int OP, typename TB, typename TO>
void launch_op(hipStream_t st, const AmgMat& M, const float* dinv, const TB* b, const float* xin,
               const float* base, float* xout, float* res, TO* out, const double* bin = nullptr,
               double* bout = nullptr, const int* done = nullptr, float* out32 = nullptr) {
    switch (M.g) {
        case 0:
            if constexpr (OP == OP_RQ)
                throw std::runtime_error("amg: a composed restriction needs a CSR layout");
            else
                launch_g<0, OP>(st, M, dinv, b, xin, base, xout, res, out, bin, bout, done, out32);
            break;
        case 4: launch_g<4, OP>(st, M, dinv, b, xin, base, xout, res, out, bin, bout, done, out32); break;
        case 8: launch_g<8, OP>(st, M, dinv, b, xin, base, xout, res, out, bin, bout, done, out32); break;
        case 16: launch_g<16, OP>(st, M, dinv, b, xin, base, xout, res, out, bin, bout, done, out32); break;
        case 32: launch_g<32, OP>(st, M, dinv, b, xin, base, xout, res, out, bin, bout, done, out32); break;
        default: launch_g<64, OP>(st, M, dinv, b, xin, base, xout, res, out, bin, bout, done, out32); break;
    }
}

struct HostSell {
    std::vector<int> sptr, col;
    std::vector<uint16_t> val;
    int nslices = 0;
};

HostSell to_sell(const mp_amg::Csr& M) {
    HostSell h;
    h.nslices = (M.nrows + SLICE - 1) / SLICE;
    h.sptr.assign(h.nslices + 1, 0);
    for (int s = 0; s < h.nslices; ++s) {
        int w = 0;
        for (int l = 0; l < SLICE; ++l) {
            const int r = s * SLICE + l;
            if (r < M.nrows) w = std::max(w, M.ptr[r + 1] - M.ptr[r]);
        }
        h.sptr[s + 1] = h.sptr[s] + sell_width4(w) * SLICE;
    }
    h.col.assign(h.sptr[h.nslices], 0);
    h.val.assign(h.sptr[h.nslices], 0);
    for (int s = 0; s < h.nslices; ++s) {
        const int w = (h.sptr[s + 1] - h.sptr[s]) / SLICE;
        for (int l = 0; l < SLICE; ++l) {
            const int r = s * SLICE + l;
            if (r >= M.nrows) continue;
            const int len = M.ptr[r + 1] - M.ptr[r];
            for (int j = 0; j < w && j < len; ++j) {
                const int slot = h.sptr[s] + sell_off(j) + l * SGROUP;
                h.col[slot] = M.col[M.ptr[r] + j];
                h.val[slot] = f2bf((float)M.val[M.ptr[r] + j]);
            }
        }
    }
    return h;
}

template <typename T>
void put(DevBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
    d.alloc(h.size());
    if (!h.empty())
        MP_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

// Uploads M in layout g (0: SELL-64x4, else CSR with g lanes per row).
void upload_mat(AmgMat& d, const mp_amg::Csr& M, int g, hipStream_t st) {
    d.n = M.nrows;
    d.g = g;
    d.nnz = M.nnz();
    if (g == 0) {
        HostSell h = to_sell(M);
        d.nslices = h.nslices;
        put(d.ptr_b, h.sptr, st);
        put(d.col_b, h.col, st);
        put(d.val_b, h.val, st);
    } else {
        d.nslices = 0;
        std::vector<uint16_t> v(M.val.size());
        for (size_t t = 0; t < v.size(); ++t) v[t] = f2bf((float)M.val[t]);
        put(d.ptr_b, M.ptr, st);
        put(d.col_b, M.col, st);
        put(d.val_b, v, st);
    }
    d.ptr = d.ptr_b.p;
    d.col = d.col_b.p;
    d.val = d.val_b.p;
}

// Reads n4 float4s (grid-stride), so that a product timed after it finds its
// matrix and vector outside the caches, as inside a V-cycle, where the finest
// level's passes evict them between two uses.
__global__ __launch_bounds__(BLOCK) void k_stream_read(long n4, const float4* __restrict__ a,
                                                       float* __restrict__ sink) {
    float acc = 0.0f;
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n4; i += (long)gridDim.x * BLOCK) {
        const float4 v = a[i];
        acc += (v.x + v.y) + (v.z + v.w);
    }
    if (acc == 1.25e-30f) sink[0] = acc;  // (the buffer holds zeros: keeps the loads only)
}

// Autotuned layout: every candidate (SELL-64x4 for rows of at most 24 entries
// on average unless its padding more than triples the stored entries; CSR
// with 4..64 lanes per row) is uploaded, then all are timed on products with
// this matrix in interleaved rounds (so clock drift hits every candidate
// alike), each product after a read of a buffer larger than the caches, each
// candidate by its median.  The set-up rule's layout stays unless another is
// more than 10% faster.
struct Flush {
    const float4* buf;
    long n4;
    float* sink;
};
int rule_layout(const mp_amg::Csr& M);
void upload_tuned(AmgMat& d, const mp_amg::Csr& M, hipStream_t st, float* xin, float* out,
                  hipEvent_t e0, hipEvent_t e1, const Flush& fl, std::string& report,
                  const char* label) {
    std::vector<int> cand;
    {
        long slots = 0;
        for (int s = 0; s * SLICE < M.nrows; ++s) {
            int w = 0;
            for (int l = 0; l < SLICE && s * SLICE + l < M.nrows; ++l)
                w = std::max(w, M.ptr[s * SLICE + l + 1] - M.ptr[s * SLICE + l]);
            slots += (long)sell_width4(w) * SLICE;
        }
        // SELL-64x4 walks a row on one lane: only for short rows (the set-up
        // rule's bound), where its padding stays small
        const double avg = (double)M.nnz() / std::max(1, M.nrows);
        if (avg <= 24.0 && slots <= 3 * std::max<long>(1, M.nnz())) cand.push_back(0);
    }
    for (int g : {4, 8, 16, 32, 64}) cand.push_back(g);
    const int nc = (int)cand.size();
    std::vector<AmgMat> trial(nc);
    for (int c = 0; c < nc; ++c) {
        upload_mat(trial[c], M, cand[c], st);
        launch_op<OP_SPMV, float, float>(st, trial[c], nullptr, (const float*)nullptr, xin,
                                         nullptr, nullptr, nullptr, out);
    }
    constexpr int rounds = 7, reps = 3;
    std::vector<std::vector<float>> t(nc);
    for (int r = 0; r < rounds; ++r)
        for (int c = 0; c < nc; ++c) {
            float tot = 0.0f;
            for (int k = 0; k < reps; ++k) {
                hipLaunchKernelGGL(k_stream_read, dim3(2048), dim3(BLOCK), 0, st, fl.n4, fl.buf,
                                   fl.sink);
                MP_HIP(hipEventRecord(e0, st));
                launch_op<OP_SPMV, float, float>(st, trial[c], nullptr, (const float*)nullptr,
                                                 xin, nullptr, nullptr, nullptr, out);
                MP_HIP(hipEventRecord(e1, st));
                MP_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                MP_HIP(hipEventElapsedTime(&ms, e0, e1));
                tot += ms;
            }
            t[c].push_back(tot / reps);
        }
    std::vector<float> med(nc);
    for (int c = 0; c < nc; ++c) {
        std::sort(t[c].begin(), t[c].end());
        med[c] = t[c][rounds / 2];
    }
    int best = 0;
    for (int c = 1; c < nc; ++c)
        if (med[c] < med[best]) best = c;
    const int rule = rule_layout(M);
    for (int c = 0; c < nc; ++c)
        if (cand[c] == rule && med[c] <= 1.1f * med[best]) best = c;
    d = std::move(trial[best]);
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%s:%s%d(%.1fus) ", label, cand[best] == 0 ? "sell" : "csr",
                  cand[best] == 0 ? 64 : cand[best], 1e3 * med[best]);
    report += buf;
}

// Rule-of-thumb layout (deterministic, used at set-up): SELL-64x4 for short
// rows, else CSR with about one lane per 5 entries of the mean row (4..64).
int rule_layout(const mp_amg::Csr& M) {
    const double avg = (double)M.nnz() / std::max(1, M.nrows);
    if (avg <= 24.0) return 0;
    int g = 4;
    while (g < 64 && g * 5 < avg) g *= 2;
    return g;
}

std::vector<float> scaled_diag_inverse(const mp_amg::Csr& A, double omega) {
    std::vector<float> d(A.nrows, 0.0f);
    for (int i = 0; i < A.nrows; ++i) {
        double a = 0.0;
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
            if (A.col[t] == i) a += A.val[t];
        d[i] = (a != 0.0) ? (float)(omega / a) : 0.0f;
    }
    return d;
}

void alloc_zero(DevBuf<float>& b, size_t n, hipStream_t st) {
    b.alloc(n);
    MP_HIP(hipMemsetAsync(b.p, 0, sizeof(float) * b.n, st));
}

// Block-diagonal union of per-problem matrices: block q's rows start at
// roff[q] and its columns at coff[q].
mp_amg::Csr block_diag(const std::vector<const mp_amg::Csr*>& m, const std::vector<int>& roff,
                       const std::vector<int>& coff, int nrows, int ncols) {
    mp_amg::Csr out;
    out.nrows = nrows;
    out.ncols = ncols;
    std::vector<int> len(nrows, 0);
    for (size_t q = 0; q < m.size(); ++q)
        for (int i = 0; i < m[q]->nrows; ++i) len[roff[q] + i] = m[q]->ptr[i + 1] - m[q]->ptr[i];
    out.ptr.assign(nrows + 1, 0);
    for (int i = 0; i < nrows; ++i) out.ptr[i + 1] = out.ptr[i] + len[i];
    out.col.resize(out.ptr[nrows]);
    out.val.resize(out.ptr[nrows]);
    for (size_t q = 0; q < m.size(); ++q)
        for (int i = 0; i < m[q]->nrows; ++i) {
            int t = out.ptr[roff[q] + i];
            for (int e = m[q]->ptr[i]; e < m[q]->ptr[i + 1]; ++e, ++t) {
                out.col[t] = m[q]->col[e] + coff[q];
                out.val[t] = m[q]->val[e];
            }
        }
    return out;
}

// X - diag(d) Y (d null: X - Y) over the union of the two patterns, columns
// sorted; xv (may be null) receives X's values in the same slots (0 where X
// has no entry).
mp_amg::Csr minus_scaled(const mp_amg::Csr& X, const std::vector<float>* d, const mp_amg::Csr& Y,
                         std::vector<double>* xv) {
    if (X.nrows != Y.nrows || X.ncols != Y.ncols) throw std::runtime_error("amg: shape mismatch");
    mp_amg::Csr out;
    out.nrows = X.nrows;
    out.ncols = X.ncols;
    out.ptr.assign(X.nrows + 1, 0);
    std::vector<int> mark(X.ncols, -1), pos(X.ncols, 0), cols;
    std::vector<double> vx, vy;
    if (xv) xv->clear();
    for (int i = 0; i < X.nrows; ++i) {
        cols.clear();
        for (const mp_amg::Csr* M : {&X, &Y})
            for (int t = M->ptr[i]; t < M->ptr[i + 1]; ++t)
                if (mark[M->col[t]] != i) {
                    mark[M->col[t]] = i;
                    cols.push_back(M->col[t]);
                }
        std::sort(cols.begin(), cols.end());
        for (size_t k = 0; k < cols.size(); ++k) pos[cols[k]] = (int)k;
        vx.assign(cols.size(), 0.0);
        vy.assign(cols.size(), 0.0);
        for (int t = X.ptr[i]; t < X.ptr[i + 1]; ++t) vx[pos[X.col[t]]] += X.val[t];
        for (int t = Y.ptr[i]; t < Y.ptr[i + 1]; ++t) vy[pos[Y.col[t]]] += Y.val[t];
        const double di = d ? (double)(*d)[i] : 1.0;
        for (size_t k = 0; k < cols.size(); ++k) {
            out.col.push_back(cols[k]);
            out.val.push_back(vx[k] - di * vy[k]);
            if (xv) xv->push_back(vx[k]);
        }
        out.ptr[i + 1] = (int)out.col.size();
    }
    return out;
}

}  // namespace

void AmgDevice::setup(hipStream_t st, const SellView& sv, const double* val64_dev,
                      const SellView& sv_apply, const float* val32_dev,
                      const std::vector<std::pair<int, int>>& blocks,
                      const std::vector<int>& bordered, const mp_amg::Options& opt) {
    const auto t0 = std::chrono::steady_clock::now();
    if (blocks.empty()) throw std::runtime_error("amg: no problems");
    if (bordered.size() != blocks.size()) throw std::runtime_error("amg: border flags");
    // finest operator of each problem as CSR in device order, local indices
    // (padding slots dropped)
    std::vector<int> hsptr(sv.nslices + 1), hlen((size_t)sv.nslices * SLICE);
    MP_HIP(hipMemcpyAsync(hsptr.data(), sv.sptr, sizeof(int) * hsptr.size(), hipMemcpyDeviceToHost, st));
    MP_HIP(hipMemcpyAsync(hlen.data(), sv.len, sizeof(int) * hlen.size(), hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
    const int slots = hsptr[sv.nslices];
    std::vector<int> hcol(slots);
    std::vector<double> hval(slots);
    MP_HIP(hipMemcpyAsync(hcol.data(), sv.col, sizeof(int) * slots, hipMemcpyDeviceToHost, st));
    MP_HIP(hipMemcpyAsync(hval.data(), val64_dev, sizeof(double) * slots, hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
    const int nb = (int)blocks.size();
    std::vector<mp_amg::Csr> A0s(nb);
    for (int q = 0; q < nb; ++q) {
        const int r0 = blocks[q].first, n = blocks[q].second;
        mp_amg::Csr& A0 = A0s[q];
        A0.nrows = A0.ncols = n;
        A0.ptr.assign(n + 1, 0);
        for (int i = 0; i < n; ++i) A0.ptr[i + 1] = A0.ptr[i] + hlen[r0 + i];
        A0.col.resize(A0.ptr[n]);
        A0.val.resize(A0.ptr[n]);
        for (int i = 0; i < n; ++i) {
            const int base = sell_base(hsptr.data(), r0 + i);
            for (int j = 0; j < hlen[r0 + i]; ++j) {
                const int slot = base + sell_off(j);
                const int c = hcol[slot] - r0;
                if (c < 0 || c >= n) throw std::runtime_error("amg: a stencil leaves its problem");
                A0.col[A0.ptr[i] + j] = c;
                A0.val[A0.ptr[i] + j] = hval[slot];
            }
        }
    }
    { std::vector<int>().swap(hcol); std::vector<double>().swap(hval); }
    // one hierarchy per problem; in a batch all of the same depth (shallower
    // ones coarsen further, to a smaller dense coarse solve).  At least two
    // levels: the bordered dense inverse alone maps onto mean-free vectors,
    // which would leave the bordered system's constraint row of the
    // preconditioned operator zero; smoothing on the finest level keeps it.
    std::vector<mp_amg::Hierarchy> hs(nb);
    int L = 2;
    for (int q = 0; q < nb; ++q) {
        hs[q] = mp_amg::build(A0s[q], opt, bordered[q] != 0);
        L = std::max(L, hs[q].levels());
    }
    for (int q = 0; q < nb; ++q)
        if (hs[q].levels() < L) {
            mp_amg::Options o = opt;
            o.coarse_max = 0;
            o.max_levels = L;
            hs[q] = mp_amg::build(A0s[q], o, bordered[q] != 0);
            if (hs[q].levels() != L)
                throw std::runtime_error("amg: problem " + std::to_string(q) +
                                         " cannot be coarsened to " + std::to_string(L) +
                                         " levels");
        }
    // row offsets per level: level 0 is the device rows, levels >= 1 stack
    // the problems' rows in problem order
    const int n0 = sv_apply.nv;
    std::vector<std::vector<int>> off(L, std::vector<int>(nb, 0));
    std::vector<int> ntot(L, 0);
    for (int q = 0; q < nb; ++q) off[0][q] = blocks[q].first;
    ntot[0] = n0;
    for (int l = 1; l < L; ++l)
        for (int q = 0; q < nb; ++q) {
            off[l][q] = ntot[l];
            ntot[l] += hs[q].A[l].nrows;
        }
    // host copies of the block-diagonal hierarchy (A[0] stays empty)
    std::vector<mp_amg::Csr> Ab(L), Pb(L > 1 ? L - 1 : 0), Rb(L > 1 ? L - 1 : 0);
    for (int l = 0; l < L; ++l) {
        std::vector<const mp_amg::Csr*> ma(nb), mp_(nb), mr(nb);
        for (int q = 0; q < nb; ++q) {
            ma[q] = l == 0 ? &A0s[q] : &hs[q].A[l];
            if (l + 1 < L) {
                mp_[q] = &hs[q].P[l];
                mr[q] = &hs[q].R[l];
            }
        }
        if (l > 0) Ab[l] = block_diag(ma, off[l], off[l], ntot[l], ntot[l]);
        if (l + 1 < L) {
            Pb[l] = block_diag(mp_, off[l], off[l + 1], ntot[l], ntot[l + 1]);
            Rb[l] = block_diag(mr, off[l + 1], off[l], ntot[l + 1], ntot[l]);
        }
    }
    lv.clear();
    lv.resize(L);
    std::vector<std::vector<float>> dvs(L);  // host copies of omega / a_ii per level
    rows.assign(L, 0);
    nnz.assign(L, 0);
    layout.clear();
    for (int l = 0; l < L; ++l) {
        AmgLevelDev& d = lv[l];
        d.n = ntot[l];
        long nz = 0;
        if (l == 0) {
            int nreal = 0;
            for (int q = 0; q < nb; ++q) {
                nz += A0s[q].nnz();
                nreal += A0s[q].nrows;
            }
            rows[0] = nreal;
        } else {
            nz = Ab[l].nnz();
            rows[l] = ntot[l];
        }
        nnz[l] = nz;
        std::vector<float> dv(ntot[l], 0.0f);  // level 0: rows outside the problems get 0
        for (int q = 0; q < nb; ++q) {
            const std::vector<float> dq =
                scaled_diag_inverse(l == 0 ? A0s[q] : hs[q].A[l], omega);
            std::copy(dq.begin(), dq.end(), dv.begin() + off[l][q]);
        }
        put(d.dinv, dv, st);
        dvs[l] = dv;
        if (l == 0) {
            d.A.n = n0;
            d.A.nslices = sv_apply.nslices;
            d.A.g = 0;
            d.A.nnz = nz;
            d.A.ptr = sv_apply.sptr;
            d.A.col = sv_apply.col;
            d.A.col16 = sv_apply.col16;
            d.A.cbase = sv_apply.cbase;
            // BF16 copy of the solver's FP32 slots (padding slots stay 0)
            d.A.val_b.alloc((size_t)slots);
            hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((slots + BLOCK - 1) / BLOCK)),
                               dim3(BLOCK), 0, st, (long)slots, val32_dev, d.A.val_b.p);
            MP_HIP(hipGetLastError());
            d.A.val = d.A.val_b.p;
        } else if (l + 1 < L) {
            upload_mat(d.A, Ab[l], rule_layout(Ab[l]), st);
        }
        const size_t vl = (size_t)ntot[l] + 64;
        alloc_zero(d.x, vl, st);
        alloc_zero(d.res, vl, st);
        if (l > 0) {
            alloc_zero(d.b, vl, st);
            alloc_zero(d.y, vl, st);
        }
        if (l + 1 < L) {
            upload_mat(d.P, Pb[l], rule_layout(Pb[l]), st);
            upload_mat(d.R, Rb[l], rule_layout(Rb[l]), st);
        }
    }
    // coarsest level: the dense inverses, one block per problem
    nc = ntot[L - 1];
    std::vector<float> g;
    std::vector<int> hfirst(nc), hsize(nc);
    std::vector<long> hdata(nc);
    for (int q = 0; q < nb; ++q) {
        const int m = hs[q].nc, first = off[L - 1][q];
        const long d0 = (long)g.size();
        for (double v : hs[q].coarse_inv) g.push_back((float)v);
        for (int r = 0; r < m; ++r) {
            hfirst[first + r] = first;
            hsize[first + r] = m;
            hdata[first + r] = d0 + (long)r * m;
        }
    }
    put(cinv, g, st);
    // composed steps below level 1 (three or more levels)
    composed = L >= 3;
    plain_cycle = false;
    hT.assign(L, mp_amg::Csr());
    if (composed) {
        for (int l = 1; l <= L - 2; ++l) {
            const mp_amg::Csr AP = mp_amg::multiply(Ab[l], Pb[l], opt.threads);
            hT[l] = minus_scaled(Pb[l], &dvs[l], AP, nullptr);
            upload_mat(lv[l].T, hT[l], rule_layout(hT[l]), st);
            if (l >= 2) {
                mp_amg::Csr DR = Rb[l - 1];
                for (int i = 0; i < DR.nrows; ++i)
                    for (int t = DR.ptr[i]; t < DR.ptr[i + 1]; ++t) DR.val[t] *= (double)dvs[l][i];
                const mp_amg::Csr ADR = mp_amg::multiply(Ab[l], DR, opt.threads);
                std::vector<double> rv;
                const mp_amg::Csr Q = minus_scaled(Rb[l - 1], nullptr, ADR, &rv);
                upload_mat(lv[l].Q, Q, std::max(4, rule_layout(Q)), st);
                std::vector<uint16_t> r16(rv.size());
                for (size_t t = 0; t < rv.size(); ++t) r16[t] = f2bf((float)rv[t]);
                put(lv[l].Q.val2_b, r16, st);
                lv[l].Q.val2 = lv[l].Q.val2_b.p;
            }
        }
        // one problem: the coarsest solve times R_{L-2}, one dense row per
        // coarsest row
        std::vector<uint16_t> cv;
        std::vector<int> cf(nc), cs(nc);
        std::vector<long> cd(nc);
        for (int q = 0; q < nb && nb == 1; ++q) {
            const int m = hs[q].nc;
            const mp_amg::Csr& Rq = hs[q].R[L - 2];
            if (Rq.nrows != m) throw std::runtime_error("amg: coarsest level size");
            const int n = Rq.ncols;
            std::vector<double> row(n);
            for (int r = 0; r < m; ++r) {
                std::fill(row.begin(), row.end(), 0.0);
                for (int k = 0; k < m; ++k) {
                    const double gk = hs[q].coarse_inv[(size_t)r * m + k];
                    for (int t = Rq.ptr[k]; t < Rq.ptr[k + 1]; ++t) row[Rq.col[t]] += gk * Rq.val[t];
                }
                const int rr = off[L - 1][q] + r;
                cf[rr] = off[L - 2][q];
                cs[rr] = n;
                cd[rr] = (long)cv.size();
                for (int j = 0; j < n; ++j) cv.push_back(f2bf((float)row[j]));
                while (cv.size() % 4) cv.push_back(0);
            }
        }
        put(cr_val, cv, st);
        put(cr_first, cf, st);
        put(cr_size, cs, st);
        put(cr_data, cd, st);
    }
    blocked = nb > 1;
    if (blocked) {
        put(cfirst, hfirst, st);
        put(csize, hsize, st);
        put(cdata, hdata, st);
    }
    MP_HIP(hipStreamSynchronize(st));
    hA = std::move(Ab);
    hP = std::move(Pb);
    hR = std::move(Rb);
    for (int l = 0; l + 1 < L; ++l) {
        char buf[96];
        std::snprintf(buf, sizeof(buf), "A%d:%d P%d:%d R%d:%d ", l, lv[l].A.g, l, lv[l].P.g, l,
                      lv[l].R.g);
        layout += buf;
        if (composed && l >= 1) {
            std::snprintf(buf, sizeof(buf), "T%d:%d ", l, lv[l].T.g);
            layout += buf;
            if (l >= 2) {
                std::snprintf(buf, sizeof(buf), "Q%d:%d ", l, lv[l].Q.g);
                layout += buf;
            }
        }
    }
    setup_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void AmgDevice::set_border_rows(hipStream_t st, const std::vector<int>& r) {
    if (r.empty()) return;
    put(border_rows, r, st);
    MP_HIP(hipStreamSynchronize(st));
}

void AmgDevice::retune(hipStream_t st) {
    const int L = (int)lv.size();
    if (L < 2) return;
    size_t big = 64;
    for (int l = 0; l < L; ++l) big = std::max(big, (size_t)lv[l].n + 64);
    DevBuf<float> tx, ty, fb;
    alloc_zero(tx, big, st);
    alloc_zero(ty, big, st);
    alloc_zero(fb, (size_t)128 << 20, st);  // 512 MiB, above the L2s and the last-level cache
    const Flush fl{reinterpret_cast<const float4*>(fb.p), (long)(fb.n / 4), ty.p + big - 1};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MP_HIP(hipEventCreate(&e0));
    MP_HIP(hipEventCreate(&e1));
    layout.clear();
    try {
        for (int l = 0; l + 1 < L; ++l) {
            AmgLevelDev& d = lv[l];
            // the composed cycle uses A_1, P_0, R_0, the T_l (the Q_l keep
            // their CSR layout) and, in a batch, R_{L-2}; the plain cycle
            // every A_l, P_l, R_l
            const bool useA = l > 0 && (!composed || l == 1);
            const bool useP = !composed || l == 0;
            const bool useR = !composed || l == 0 || (blocked && l == L - 2);
            if (useA)
                upload_tuned(d.A, hA[l], st, tx.p, ty.p, e0, e1, fl, layout,
                             ("A" + std::to_string(l)).c_str());
            if (useP)
                upload_tuned(d.P, hP[l], st, tx.p, ty.p, e0, e1, fl, layout,
                             ("P" + std::to_string(l)).c_str());
            if (useR)
                upload_tuned(d.R, hR[l], st, tx.p, ty.p, e0, e1, fl, layout,
                             ("R" + std::to_string(l)).c_str());
            if (composed && l >= 1)
                upload_tuned(d.T, hT[l], st, tx.p, ty.p, e0, e1, fl, layout,
                             ("T" + std::to_string(l)).c_str());
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    MP_HIP(hipEventDestroy(e0));
    MP_HIP(hipEventDestroy(e1));
    MP_HIP(hipStreamSynchronize(st));
}

void AmgDevice::apply(hipStream_t st, const double* in, double* out, int bordered,
                      const int* done, bool prescaled, float* out32) {
    const int nv = lv[0].n;
    // one problem: its border unknown (row nv) passes through in the kernels;
    // a batch copies every border unknown at the end
    const double* bin = (bordered && !blocked) ? in + nv : nullptr;
    double* bout = (bordered && !blocked) ? out + nv : nullptr;
    vcycle(st, in, out, bin, bout, done, prescaled, out32);
    if (bordered && blocked && border_rows.n > 0)
        hipLaunchKernelGGL(k_border_copy, dim3((border_rows.n + 63) / 64), dim3(64), 0, st,
                           (int)border_rows.n, border_rows.p, in, out);
    MP_HIP(hipGetLastError());
}

template <typename TI>
void AmgDevice::apply_basis(hipStream_t st, const TI* in, float* out, const int* done) {
    vcycle(st, in, out, nullptr, nullptr, done, true, nullptr);
    MP_HIP(hipGetLastError());
}
template void AmgDevice::apply_basis<float>(hipStream_t, const float*, float*, const int*);
template void AmgDevice::apply_basis<double>(hipStream_t, const double*, float*, const int*);

template <typename TI, typename TO>
void AmgDevice::vcycle(hipStream_t st, const TI* in, TO* out, const double* bin, double* bout,
                       const int* done, bool prescaled, float* out32) {
    const int L = (int)lv.size();
    auto coarse = [&](const auto* b, auto* y, float* y32, const double* cbin, double* cbout) {
        using TB = std::remove_cv_t<std::remove_pointer_t<decltype(b)>>;
        using TY = std::remove_pointer_t<decltype(y)>;
        if (blocked)
            hipLaunchKernelGGL((k_amg_blockdense<TB, TY>), dim3((nc + 3) / 4), dim3(BLOCK), 0, st,
                               nc, cinv.p, cfirst.p, csize.p, cdata.p, b, y, y32);
        else
            hipLaunchKernelGGL((k_amg_dense<TB, TY>), dim3((nc + 3) / 4), dim3(BLOCK), 0, st, nc,
                               cinv.p, b, y, cbin, cbout, done, y32);
    };
    {  // (setup builds at least two levels)
        // downward: level 0 from the input (FP64, or an FP32 basis vector)
        if (prescaled)
            launch_op<OP_PRES, TI, float>(st, lv[0].A, nullptr, in, lv[0].x.p, nullptr,
                                          nullptr, lv[0].res.p, (float*)nullptr, bin, bout,
                                          done);
        else
            launch_op<OP_PRE, TI, float>(st, lv[0].A, lv[0].dinv.p, in, nullptr, nullptr,
                                             lv[0].x.p, lv[0].res.p, (float*)nullptr, bin, bout,
                                             done);
        if (composed && !plain_cycle) {
            // level 0 -> 1 and level 1's pre-smoothing as below; then one
            // composed product per coarse level down, the composed coarsest
            // solve, one composed product per level up to level 1
            launch_op<OP_SPMV, float, float>(st, lv[0].R, nullptr, (const float*)nullptr,
                                             lv[0].res.p, nullptr, nullptr, nullptr, lv[1].b.p,
                                             nullptr, nullptr, done);
            launch_op<OP_PRE, float, float>(st, lv[1].A, lv[1].dinv.p, lv[1].b.p, nullptr,
                                            nullptr, lv[1].x.p, lv[1].res.p, (float*)nullptr,
                                            nullptr, nullptr, done);
            for (int l = 2; l <= L - 2; ++l)
                launch_op<OP_RQ, float, float>(st, lv[l].Q, nullptr, (const float*)nullptr,
                                               lv[l - 1].res.p, nullptr, lv[l].b.p, lv[l].res.p,
                                               (float*)nullptr, nullptr, nullptr, done);
            if (!blocked) {
                hipLaunchKernelGGL(k_amg_rowdense, dim3(nc), dim3(RD_THREADS), 0, st, nc,
                                   cr_val.p, cr_first.p, cr_size.p, cr_data.p,
                                   (const float*)lv[L - 2].res.p, lv[L - 1].y.p, done);
            } else {
                launch_op<OP_SPMV, float, float>(st, lv[L - 2].R, nullptr, (const float*)nullptr,
                                                 lv[L - 2].res.p, nullptr, nullptr, nullptr,
                                                 lv[L - 1].b.p, nullptr, nullptr, done);
                coarse((const float*)lv[L - 1].b.p, lv[L - 1].y.p, (float*)nullptr,
                       (const double*)nullptr, (double*)nullptr);
            }
            for (int l = L - 2; l >= 1; --l)
                launch_op<OP_TPOST, float, float>(st, lv[l].T, lv[l].dinv.p, lv[l].b.p,
                                                  lv[l + 1].y.p, lv[l].res.p, nullptr, nullptr,
                                                  lv[l].y.p, nullptr, nullptr, done);
            AmgLevelDev& f = lv[0];
            launch_op<OP_SPMV, float, float>(st, f.P, nullptr, (const float*)nullptr, lv[1].y.p,
                                             f.x.p, nullptr, nullptr, f.x.p, nullptr, nullptr,
                                             done);
            if (post_timer) post_timer->begin(st);
            launch_op<OP_POST, TI, TO>(st, f.A, f.dinv.p, in, f.x.p, nullptr, nullptr, nullptr,
                                       out, nullptr, nullptr, done, out32);
            if (post_timer) post_timer->end(st);
            return;
        }
        for (int l = 0; l + 1 < L; ++l) {
            AmgLevelDev& f = lv[l];
            AmgLevelDev& c = lv[l + 1];
            launch_op<OP_SPMV, float, float>(st, f.R, nullptr, (const float*)nullptr, f.res.p,
                                             nullptr, nullptr, nullptr, c.b.p, nullptr, nullptr,
                                             done);
            if (l + 2 == L) {  // (the border unknown was copied at level 0)
                coarse((const float*)c.b.p, c.y.p, (float*)nullptr, (const double*)nullptr,
                       (double*)nullptr);
            } else {
                launch_op<OP_PRE, float, float>(st, c.A, c.dinv.p, c.b.p, nullptr, nullptr, c.x.p,
                                                c.res.p, (float*)nullptr, nullptr, nullptr, done);
            }
        }
        // upward
        for (int l = L - 2; l >= 0; --l) {
            AmgLevelDev& f = lv[l];
            AmgLevelDev& c = lv[l + 1];
            launch_op<OP_SPMV, float, float>(st, f.P, nullptr, (const float*)nullptr, c.y.p, f.x.p,
                                             nullptr, nullptr, f.x.p, nullptr, nullptr, done);
            if (l == 0) {
                if (post_timer) post_timer->begin(st);
                launch_op<OP_POST, TI, TO>(st, f.A, f.dinv.p, in, f.x.p, nullptr, nullptr,
                                           nullptr, out, nullptr, nullptr, done, out32);
                if (post_timer) post_timer->end(st);
            }
            else
                launch_op<OP_POST, float, float>(st, f.A, f.dinv.p, f.b.p, f.x.p, nullptr,
                                                 nullptr, nullptr, f.y.p, nullptr, nullptr, done);
        }
    }
}

}  // namespace mp
