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
// through unchanged.  A matrix is SELL-64x4 (one wavefront per slice of 64
// rows, four neighbours of a row side by side, padding slots with value 0)
// or CSR with several lanes per row, whichever suits its row lengths.
#include "amg.hpp"

#include <chrono>
#include <cstring>

namespace mp {

namespace {

enum { OP_PRE = 0, OP_SPMV = 1, OP_POST = 2, OP_PRES = 3 };

__host__ __device__ inline int nblk_rows(int nslices) { return (nslices * SLICE + BLOCK - 1) / BLOCK; }

__device__ __forceinline__ bool stopped(const int* done) { return done && *done; }

struct MatView {
    int n, nslices;
    const int* ptr;
    const int* col;
    const uint16_t* val;    // BF16
    const uint16_t* col16;  // compressed columns (SELL-64x4 only; may be null)
    const int* cbase;
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
    float acc = 0.0f;
    int i;
    if constexpr (G == 0) {
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
            }
            for (; t < t1; t += G)
                acc = fmaf(bf2f(m.val[t]), gathered<OP>(m.col[t], b, xin, dinv), acc);
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
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

template <int G, int OP, typename TB, typename TO>
void launch_g(hipStream_t st, const AmgMat& M, const float* dinv, const TB* b, const float* xin,
              const float* base, float* xout, float* res, TO* out, const double* bin,
              double* bout, const int* done, float* out32) {
    const MatView v{M.n, M.nslices, M.ptr, M.col, M.val, M.col16, M.cbase};
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
        case 0: launch_g<0, OP>(st, M, dinv, b, xin, base, xout, res, out, bin, bout, done, out32); break;
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

// Autotuned layout: every candidate (SELL-64x4 unless its padding more than
// triples the stored entries; CSR with 4..64 lanes per row) is uploaded and
// timed on a product with this matrix; the fastest stays.
void upload_tuned(AmgMat& d, const mp_amg::Csr& M, hipStream_t st, float* xin, float* out,
                  hipEvent_t e0, hipEvent_t e1, std::string& report, const char* label) {
    std::vector<int> cand;
    {
        long slots = 0;
        for (int s = 0; s * SLICE < M.nrows; ++s) {
            int w = 0;
            for (int l = 0; l < SLICE && s * SLICE + l < M.nrows; ++l)
                w = std::max(w, M.ptr[s * SLICE + l + 1] - M.ptr[s * SLICE + l]);
            slots += (long)sell_width4(w) * SLICE;
        }
        if (slots <= 3 * std::max<long>(1, M.nnz())) cand.push_back(0);
    }
    for (int g : {4, 8, 16, 32, 64}) cand.push_back(g);
    float best_ms = 1e30f;
    int best = cand.front();
    for (int g : cand) {
        AmgMat trial;
        upload_mat(trial, M, g, st);
        launch_op<OP_SPMV, float, float>(st, trial, nullptr, (const float*)nullptr, xin, nullptr,
                                         nullptr, nullptr, out);
        MP_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 5; ++r)
            launch_op<OP_SPMV, float, float>(st, trial, nullptr, (const float*)nullptr, xin,
                                             nullptr, nullptr, nullptr, out);
        MP_HIP(hipEventRecord(e1, st));
        MP_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        MP_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = g;
        }
    }
    upload_mat(d, M, best, st);
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%s%s:%s%d(%.1fus) ", report.empty() ? "" : "", label,
                  best == 0 ? "sell" : "csr", best == 0 ? 64 : best, 1e3 * best_ms / 5);
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

}  // namespace

void AmgDevice::setup(hipStream_t st, const SellView& sv, const double* val64_dev,
                      const SellView& sv_apply, const float* val32_dev,
                      const mp_amg::Options& opt) {
    const auto t0 = std::chrono::steady_clock::now();
    // finest operator as CSR in device order (padding slots dropped)
    const int nv = sv.nv;
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
    mp_amg::Csr A0;
    A0.nrows = A0.ncols = nv;
    A0.ptr.assign(nv + 1, 0);
    for (int i = 0; i < nv; ++i) A0.ptr[i + 1] = A0.ptr[i] + hlen[i];
    A0.col.resize(A0.ptr[nv]);
    A0.val.resize(A0.ptr[nv]);
    for (int i = 0; i < nv; ++i) {
        const int base = sell_base(hsptr.data(), i);
        for (int j = 0; j < hlen[i]; ++j) {
            const int slot = base + sell_off(j);
            A0.col[A0.ptr[i] + j] = hcol[slot];
            A0.val[A0.ptr[i] + j] = hval[slot];
        }
    }
    { std::vector<int>().swap(hcol); std::vector<double>().swap(hval); }
    mp_amg::Hierarchy h = mp_amg::build(A0, opt);
    const int L = h.levels();
    lv.clear();
    lv.resize(L);
    rows.assign(L, 0);
    nnz.assign(L, 0);
    layout.clear();
    for (int l = 0; l < L; ++l) {
            const mp_amg::Csr& A = (l == 0) ? A0 : h.A[l];
            AmgLevelDev& d = lv[l];
            d.n = A.nrows;
            rows[l] = A.nrows;
            nnz[l] = A.nnz();
            put(d.dinv, scaled_diag_inverse(A, omega), st);
            if (l == 0) {
                d.A.n = nv;
                d.A.nslices = sv_apply.nslices;
                d.A.g = 0;
                d.A.nnz = A.nnz();
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
                upload_mat(d.A, A, rule_layout(A), st);
            }
            const size_t vl = (size_t)A.nrows + 64;
            alloc_zero(d.x, vl, st);
            alloc_zero(d.res, vl, st);
            if (l > 0) {
                alloc_zero(d.b, vl, st);
                alloc_zero(d.y, vl, st);
            }
            if (l + 1 < L) {
                upload_mat(d.P, h.P[l], rule_layout(h.P[l]), st);
                upload_mat(d.R, h.R[l], rule_layout(h.R[l]), st);
            }
    }
    nc = h.nc;
    std::vector<float> g(h.coarse_inv.size());
    for (size_t t = 0; t < g.size(); ++t) g[t] = (float)h.coarse_inv[t];
    put(cinv, g, st);
    MP_HIP(hipStreamSynchronize(st));
    hA = std::move(h.A);
    hP = std::move(h.P);
    hR = std::move(h.R);
    for (int l = 0; l + 1 < L; ++l) {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "A%d:%d P%d:%d R%d:%d ", l, lv[l].A.g, l, lv[l].P.g, l,
                      lv[l].R.g);
        layout += buf;
    }
    setup_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void AmgDevice::retune(hipStream_t st) {
    const int L = (int)lv.size();
    if (L < 2) return;
    size_t big = 64;
    for (int l = 0; l < L; ++l) big = std::max(big, (size_t)lv[l].n + 64);
    DevBuf<float> tx, ty;
    alloc_zero(tx, big, st);
    alloc_zero(ty, big, st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MP_HIP(hipEventCreate(&e0));
    MP_HIP(hipEventCreate(&e1));
    layout.clear();
    try {
        for (int l = 0; l + 1 < L; ++l) {
            AmgLevelDev& d = lv[l];
            if (l > 0) upload_tuned(d.A, hA[l], st, tx.p, ty.p, e0, e1, layout,
                                    ("A" + std::to_string(l)).c_str());
            upload_tuned(d.P, hP[l], st, tx.p, ty.p, e0, e1, layout,
                         ("P" + std::to_string(l)).c_str());
            upload_tuned(d.R, hR[l], st, tx.p, ty.p, e0, e1, layout,
                         ("R" + std::to_string(l)).c_str());
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
    const int L = (int)lv.size();
    const int nv = lv[0].n;
    const double* bin = bordered ? in + nv : nullptr;
    double* bout = bordered ? out + nv : nullptr;
    if (L == 1) {
        hipLaunchKernelGGL((k_amg_dense<double, double>), dim3((nc + 3) / 4), dim3(BLOCK), 0, st,
                           nc, cinv.p, in, out, bin, bout, done, out32);
        MP_HIP(hipGetLastError());
        return;
    }
    // downward: level 0 from the FP64 input
    if (prescaled)
        launch_op<OP_PRES, double, float>(st, lv[0].A, nullptr, in, lv[0].x.p, nullptr, nullptr,
                                          lv[0].res.p, (float*)nullptr, bin, bout, done);
    else
        launch_op<OP_PRE, double, float>(st, lv[0].A, lv[0].dinv.p, in, nullptr, nullptr,
                                         lv[0].x.p, lv[0].res.p, (float*)nullptr, bin, bout, done);
    for (int l = 0; l + 1 < L; ++l) {
        AmgLevelDev& f = lv[l];
        AmgLevelDev& c = lv[l + 1];
        launch_op<OP_SPMV, float, float>(st, f.R, nullptr, (const float*)nullptr, f.res.p, nullptr,
                                         nullptr, nullptr, c.b.p, nullptr, nullptr, done);
        if (l + 2 == L)
            hipLaunchKernelGGL((k_amg_dense<float, float>), dim3((nc + 3) / 4), dim3(BLOCK), 0, st,
                               nc, cinv.p, c.b.p, c.y.p, (const double*)nullptr, (double*)nullptr,
                               done, (float*)nullptr);
        else
            launch_op<OP_PRE, float, float>(st, c.A, c.dinv.p, c.b.p, nullptr, nullptr, c.x.p,
                                            c.res.p, (float*)nullptr, nullptr, nullptr, done);
    }
    // upward
    for (int l = L - 2; l >= 0; --l) {
        AmgLevelDev& f = lv[l];
        AmgLevelDev& c = lv[l + 1];
        launch_op<OP_SPMV, float, float>(st, f.P, nullptr, (const float*)nullptr, c.y.p, f.x.p,
                                         nullptr, nullptr, f.x.p, nullptr, nullptr, done);
        if (l == 0)
            launch_op<OP_POST, double, double>(st, f.A, f.dinv.p, in, f.x.p, nullptr, nullptr,
                                               nullptr, out, nullptr, nullptr, done, out32);
        else
            launch_op<OP_POST, float, float>(st, f.A, f.dinv.p, f.b.p, f.x.p, nullptr, nullptr,
                                             nullptr, f.y.p, nullptr, nullptr, done);
    }
    MP_HIP(hipGetLastError());
}

}  // namespace mp
