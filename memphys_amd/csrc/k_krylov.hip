// Part (c): pressure-Poisson Krylov solve on the RBF-FD operator.
//
// BiCGSTAB as Eigen 3.3.7's BiCGSTAB.h (the reference's eigen_ilu_bicgstab
// branch, solver.cpp:337-350), right-preconditioned by Jacobi.  The Jacobi
// scaling is folded into the stored operator (A D^-1, column scaling) so the
// preconditioner costs nothing per iteration and x = D^-1 x~ at the end.
// With the multigrid preconditioner the stored operator is A (scaling vector
// of ones) and p, s pass through one V-cycle each (k_amg.hip) before their
// products, as in Eigen's preconditioned loop (y = M^-1 p, z = M^-1 s).
//
// Batches.  One launch serves every independent system of a batch (the
// block-diagonal union of several point clouds): each system owns whole row
// blocks, carries its own scalars (alpha, omega, convergence flag ...) and
// stops on its own; a block looks up its system once and returns at entry
// when that system is done.  A single system is the one-system batch.
//
// Per iteration: p-update, SpMV + partial r0.v, finish(alpha), s-update,
// SpMV + partial (t.s, t.t), finish(omega), x/r-update + partial (r.r, r0.r),
// finish(loop head).  Reductions are two-stage and deterministic: every block
// stores its partial sums (plain stores, no same-address atomics), then one
// 1024-thread block per system sums that system's partials in block order and
// updates its scalars in device memory.  The host launches iterations in
// chunks (replayed from a hipGraph) and polls the flags asynchronously.
//
// The full-Neumann border (solver.cpp:205-213) is applied as a rank-one
// term: the rows of a system add x[b] (b = row0 + nv, its border unknown);
// row b is sum_j D^-1_j x_j over the system's rows, formed by the finishing
// kernel.
#include "krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mp {

namespace {

constexpr int FIN = 1024;  // threads of the finishing kernels (one block per system)

// Block sums of NV values stored at partial[q * nb + lb] (lb: logical block)
// (plain stores: no atomics, the finishing kernel runs after the grid).
template <int NV>
__device__ __forceinline__ void block_store(double (&v)[NV], double* __restrict__ partial, int lb,
                                            int nb) {
    __shared__ double sh[NV][BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[q][wid] = v[q];
    __syncthreads();
    if (threadIdx.x < NV) {
        const int q = threadIdx.x;
        double s = sh[q][0];
        for (int t = 1; t < BLOCK / 64; ++t) s += sh[q][t];
        partial[(size_t)q * nb + lb] = s;
    }
}

// Sum of the partials of blocks [b0, b0 + cnt) per value (stride nb), in a
// fixed order; every thread of the block receives the totals.
template <int NV>
__device__ __forceinline__ void finish_sum(const double* __restrict__ partial, int nb, int b0,
                                           int cnt, double (&tot)[NV]) {
    __shared__ double sh[NV][FIN / 64];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int b = threadIdx.x; b < cnt; b += FIN)
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] += partial[(size_t)q * nb + b0 + b];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q)
        for (int off = 32; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[q][wid] = acc[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = sh[q][0];
        for (int t = 1; t < FIN / 64; ++t) s += sh[q][t];
        tot[q] = s;
    }
}

// Loop head of Eigen's BiCGSTAB: convergence / iteration-cap test, then
// rho_old = rho, rho = r0.r, restart test, beta.  A non-finite residual or a
// zero omega (which would make beta infinite) stops with done = 5.
__device__ void loop_top(KScalars* sc, double rr, double r0r) {
    sc->rr = rr;
    if (!(rr <= 1.0e300)) {
        sc->done = 5;
        return;
    }
    if (rr <= sc->tol2) {
        sc->done = 1;
        return;
    }
    if (sc->iter >= sc->maxit) {
        sc->done = 2;
        return;
    }
    if (sc->omega == 0.0) {
        sc->done = 5;
        return;
    }
    sc->rho_old = sc->rho;
    sc->rho = r0r;
    if (fabs(sc->rho) < sc->eps2 * sc->r0sq) {
        sc->done = 3;
        return;
    }
    sc->beta = (sc->rho / sc->rho_old) * (sc->alpha / sc->omega);
}

// Row of the SELL-64x4 product A x for slice s, lane = slice row: the slice
// width is uniform across the wavefront, so the neighbour loop runs in
// lockstep.  Padding slots hold a zero value and a valid column, which leaves
// the fma chain unchanged.
template <typename T, typename TM, typename TX>
__device__ __forceinline__ T sell_dot(const SellView& sv, const TM* __restrict__ a,
                                      const TX* __restrict__ x, int s, int lane) {
    // FP64 values (stencil order, 32-bit columns): two groups per iteration
    // keep the register count at eight waves per SIMD, and decoding 16-bit
    // columns there would make the compiler serialise the gathers.  FP32
    // values (column-sorted order, whose gathers are cheaper): four groups
    // and the 16-bit columns of the view.  x may be an FP32 copy of a vector
    // whose values are FP32 numbers (the V-cycle's output): same sums.
    constexpr bool wide = sizeof(TM) == 8;
    return sell4_accumulate<(wide ? 2 : 4), T>(
        sv.sptr, sv.col, wide ? nullptr : sv.col16, sv.cbase, a, s, lane,
        [](TM v) { return (T)v; }, [x](int c) { return (T)x[c]; });
}

// Slice of this wavefront under the XCD-aware numbering (lb: logical block).
__device__ __forceinline__ int wave_slice(int lb) {
    return __builtin_amdgcn_readfirstlane(lb * (BLOCK / SLICE) + (int)(threadIdx.x >> 6));
}

// System of logical block lb (block-uniform scalar load).
__device__ __forceinline__ int block_seg(const SegView& sg, int lb) {
    return __builtin_amdgcn_readfirstlane(sg.bseg[lb]);
}

// r = rhs - A~ x; r0 = r; p = v = 0 on a fresh start (mode 0, every system
// taking part); mode 1 restarts the systems flagged with done == 3.
template <typename TM, typename T>
__global__ __launch_bounds__(BLOCK) void k_residual_init(
    SellView sv, SegView sg, const TM* __restrict__ a, const T* __restrict__ dinv,
    const T* __restrict__ rhs, const T* __restrict__ x, T* __restrict__ r, T* __restrict__ r0,
    T* __restrict__ p, T* __restrict__ v, double* __restrict__ partial, const KScalars* sc,
    int mode) {
    const int nbk = sg.nrows / BLOCK;
    const int lb = xcd_block(nbk);
    if (lb < 0) return;
    const int q = block_seg(sg, lb);
    const int dn = sc[q].done;
    if (mode == 0 ? dn != 0 : dn != 3) return;
    const Seg S = sg.seg[q];
    const int iend = S.row0 + S.nv;
    const int i = lb * BLOCK + threadIdx.x;
    const int s = wave_slice(lb);
    double vals[3] = {0.0, 0.0, 0.0};  // r.r, rhs.rhs, sum D^-1 x
    if (s < sv.nslices) {
        T ax = sell_dot<T>(sv, a, x, s, threadIdx.x & 63);
        if (i < iend) {
            if (S.border) ax = ax + x[iend];
            const T ri = rhs[i] - ax;
            r[i] = ri;
            r0[i] = ri;
            if (mode == 0) {
                p[i] = (T)0;
                v[i] = (T)0;
            }
            vals[0] = (double)ri * (double)ri;
            vals[1] = (double)rhs[i] * (double)rhs[i];
            if (S.border) vals[2] = (double)dinv[i] * (double)x[i];
        }
    }
    block_store<3>(vals, partial, lb, nbk);
}

template <typename T>
__global__ __launch_bounds__(FIN) void k_finish_init(SegView sg, const T* __restrict__ rhs,
                                                     T* __restrict__ r, T* __restrict__ r0,
                                                     T* __restrict__ p, T* __restrict__ v,
                                                     const double* __restrict__ partial,
                                                     KScalars* scs, int mode, int rel_to_rhs) {
    KScalars* sc = scs + blockIdx.x;
    const int dn = sc->done;
    if (mode == 0 ? dn != 0 : dn != 3) return;
    const Seg S = sg.seg[blockIdx.x];
    double tot[3];
    finish_sum<3>(partial, sg.nrows / BLOCK, S.blk0, S.nblk, tot);
    if (threadIdx.x != 0) return;
    double rr = tot[0], bb = tot[1];
    if (S.border) {
        const int nb_ = S.row0 + S.nv;
        const T rn = rhs[nb_] - (T)tot[2];
        r[nb_] = rn;
        r0[nb_] = rn;
        if (mode == 0) {
            p[nb_] = (T)0;
            v[nb_] = (T)0;
        }
        rr += (double)rn * (double)rn;
        bb += (double)rhs[nb_] * (double)rhs[nb_];
    }
    if (mode == 0) {
        sc->r0sq = rr;
        sc->bb = bb;
        sc->tol2 = sc->tol_rel * sc->tol_rel * (rel_to_rhs ? bb : rr);
        sc->eps2 = 4.930380657631324e-32;  // Eigen: epsilon^2 of double
        sc->rho = 1.0;
        sc->alpha = 1.0;
        sc->omega = 1.0;
        sc->iter = 0;
        sc->restarts = 0;
        sc->red[0] = rr;
        sc->red[1] = bb;
        if (rel_to_rhs && bb == 0.0) {
            sc->done = 4;  // zero right-hand side: x = 0
            return;
        }
        loop_top(sc, rr, rr);
    } else {
        // restart with r0 = r = rhs - A x (Eigen: the first restart also resets
        // the iteration count)
        sc->r0sq = rr;
        sc->rho = rr;
        sc->rr = rr;
        sc->beta = (sc->rho / sc->rho_old) * (sc->alpha / sc->omega);
        if (sc->restarts++ == 0) sc->iter = 0;
        sc->done = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_update_p(SegView sg, const T* __restrict__ r,
                                                    T* __restrict__ p, const T* __restrict__ v,
                                                    const KScalars* scs,
                                                    const float* __restrict__ d0,
                                                    float* __restrict__ xs) {
    const int q = block_seg(sg, blockIdx.x);
    const KScalars* sc = scs + q;
    if (sc->done) return;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const T beta = (T)sc->beta, w = (T)sc->omega;
    const T pi = r[i] + beta * (p[i] - w * v[i]);
    p[i] = pi;
    if (xs && i < sg.seg[q].row0 + sg.seg[q].nv) xs[i] = d0[i] * (float)pi;  // V-cycle input
}

// v = A~ y (y = p, or M^-1 p with the multigrid preconditioner); partial sums
// of r0.v and of D^-1 y (border row).  yg: y's node rows, gathered (y itself
// or its FP32 copy); y: the border unknown.
template <typename TM, typename T, typename TX>
__global__ __launch_bounds__(BLOCK) void k_spmv_alpha(SellView sv, SegView sg,
                                                      const TM* __restrict__ a,
                                                      const T* __restrict__ dinv,
                                                      const T* __restrict__ y,
                                                      const TX* __restrict__ yg, T* __restrict__ v,
                                                      const T* __restrict__ r0,
                                                      double* __restrict__ partial,
                                                      const KScalars* sc) {
    const int nbk = sg.nrows / BLOCK;
    const int lb = xcd_block(nbk);
    if (lb < 0) return;
    const int q = block_seg(sg, lb);
    if (sc[q].done) return;
    const Seg S = sg.seg[q];
    const int iend = S.row0 + S.nv;
    const int i = lb * BLOCK + threadIdx.x;
    const int s = wave_slice(lb);
    double vals[2] = {0.0, 0.0};
    if (s < sv.nslices) {
        T yv = sell_dot<T>(sv, a, yg, s, threadIdx.x & 63);
        if (i < iend) {
            if (S.border) yv = yv + y[iend];
            v[i] = yv;
            vals[0] = (double)r0[i] * (double)yv;
            if (S.border) vals[1] = (double)dinv[i] * (double)yg[i];
        }
    }
    block_store<2>(vals, partial, lb, nbk);
}

template <typename T>
__global__ __launch_bounds__(FIN) void k_finish_alpha(SegView sg, T* __restrict__ v,
                                                      const T* __restrict__ r0,
                                                      const double* __restrict__ partial,
                                                      KScalars* scs) {
    KScalars* sc = scs + blockIdx.x;
    if (sc->done) return;
    const Seg S = sg.seg[blockIdx.x];
    double tot[2];
    finish_sum<2>(partial, sg.nrows / BLOCK, S.blk0, S.nblk, tot);
    if (threadIdx.x != 0) return;
    double r0v = tot[0];
    if (S.border) {
        const int nb_ = S.row0 + S.nv;
        const T yn = (T)tot[1];
        v[nb_] = yn;
        r0v += (double)r0[nb_] * (double)yn;
    }
    sc->alpha = sc->rho / r0v;
    if (!(fabs(sc->alpha) <= 1.0e300)) sc->done = 5;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_update_s(SegView sg, const T* __restrict__ r,
                                                    const T* __restrict__ v, T* __restrict__ s,
                                                    const KScalars* scs,
                                                    const float* __restrict__ d0,
                                                    float* __restrict__ xs) {
    const int q = block_seg(sg, blockIdx.x);
    const KScalars* sc = scs + q;
    if (sc->done) return;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const T si = r[i] - (T)sc->alpha * v[i];
    s[i] = si;
    if (xs && i < sg.seg[q].row0 + sg.seg[q].nv) xs[i] = d0[i] * (float)si;  // V-cycle input
}

// t = A~ z (z = s, or M^-1 s with the multigrid preconditioner); partial sums
// of t.s, t.t and D^-1 z
template <typename TM, typename T, typename TX>
__global__ __launch_bounds__(BLOCK) void k_spmv_omega(SellView sv, SegView sg,
                                                      const TM* __restrict__ a,
                                                      const T* __restrict__ dinv,
                                                      const T* __restrict__ z,
                                                      const TX* __restrict__ zg,
                                                      const T* __restrict__ s, T* __restrict__ t,
                                                      double* __restrict__ partial,
                                                      const KScalars* sc) {
    const int nbk = sg.nrows / BLOCK;
    const int lb = xcd_block(nbk);
    if (lb < 0) return;
    const int q = block_seg(sg, lb);
    if (sc[q].done) return;
    const Seg S = sg.seg[q];
    const int iend = S.row0 + S.nv;
    const int i = lb * BLOCK + threadIdx.x;
    const int sl = wave_slice(lb);
    double vals[3] = {0.0, 0.0, 0.0};
    if (sl < sv.nslices) {
        T y = sell_dot<T>(sv, a, zg, sl, threadIdx.x & 63);
        if (i < iend) {
            if (S.border) y = y + z[iend];
            t[i] = y;
            vals[0] = (double)y * (double)s[i];
            vals[1] = (double)y * (double)y;
            if (S.border) vals[2] = (double)dinv[i] * (double)zg[i];
        }
    }
    block_store<3>(vals, partial, lb, nbk);
}

template <typename T>
__global__ __launch_bounds__(FIN) void k_finish_omega(SegView sg, T* __restrict__ t,
                                                      const T* __restrict__ s,
                                                      const double* __restrict__ partial,
                                                      KScalars* scs) {
    KScalars* sc = scs + blockIdx.x;
    if (sc->done) return;
    const Seg S = sg.seg[blockIdx.x];
    double tot[3];
    finish_sum<3>(partial, sg.nrows / BLOCK, S.blk0, S.nblk, tot);
    if (threadIdx.x != 0) return;
    double ts = tot[0], tt = tot[1];
    if (S.border) {
        const int nb_ = S.row0 + S.nv;
        const T yn = (T)tot[2];
        t[nb_] = yn;
        ts += (double)yn * (double)s[nb_];
        tt += (double)yn * (double)yn;
    }
    sc->omega = (tt > 0.0) ? ts / tt : 0.0;
}

// x += alpha y + omega z (y = p, z = s without the multigrid preconditioner);
// r = s - omega t; partial sums of r.r and r0.r
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_update_xr(SegView sg, T* __restrict__ x,
                                                     T* __restrict__ r, const T* __restrict__ y,
                                                     const T* __restrict__ z,
                                                     const T* __restrict__ s,
                                                     const T* __restrict__ t,
                                                     const T* __restrict__ r0,
                                                     double* __restrict__ partial,
                                                     const KScalars* scs) {
    const KScalars* sc = scs + block_seg(sg, blockIdx.x);
    if (sc->done) return;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const T al = (T)sc->alpha, w = (T)sc->omega;
    x[i] = x[i] + (al * y[i] + w * z[i]);
    const T ri = s[i] - w * t[i];
    r[i] = ri;
    double vals[2] = {(double)ri * (double)ri, (double)r0[i] * (double)ri};
    block_store<2>(vals, partial, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(FIN) void k_finish_xr(SegView sg, const double* __restrict__ partial,
                                                   KScalars* scs) {
    KScalars* sc = scs + blockIdx.x;
    if (sc->done) return;
    const Seg S = sg.seg[blockIdx.x];
    double tot[2];
    finish_sum<2>(partial, sg.nrows / BLOCK, S.blk0, S.nblk, tot);
    if (threadIdx.x != 0) return;
    sc->iter += 1;
    loop_top(sc, tot[0], tot[1]);
}

// out = b - A~ x (FP64) for every system; partial sums of out.out, b.b, D^-1 x
__global__ __launch_bounds__(BLOCK) void k_residual64(SellView sv, SegView sg,
                                                      const double* __restrict__ a,
                                                      const double* __restrict__ dinv,
                                                      const double* __restrict__ b,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ out,
                                                      double* __restrict__ partial) {
    const int nbk = sg.nrows / BLOCK;
    const int lb = xcd_block(nbk);
    if (lb < 0) return;
    const Seg S = sg.seg[block_seg(sg, lb)];
    const int iend = S.row0 + S.nv;
    const int i = lb * BLOCK + threadIdx.x;
    const int s = wave_slice(lb);
    double vals[3] = {0.0, 0.0, 0.0};
    if (s < sv.nslices) {
        double ax = sell_dot<double>(sv, a, x, s, threadIdx.x & 63);
        if (i < iend) {
            if (S.border) ax = ax + x[iend];
            const double ri = b[i] - ax;
            out[i] = ri;
            vals[0] = ri * ri;
            vals[1] = b[i] * b[i];
            if (S.border) vals[2] = dinv[i] * x[i];
        }
    }
    block_store<3>(vals, partial, lb, nbk);
}

// (red[2]: the border unknown as the solve returns it, D^-1 x~, so the host
// reads it with the residual)
__global__ __launch_bounds__(FIN) void k_finish_res64(SegView sg, const double* __restrict__ b,
                                                      double* __restrict__ out,
                                                      const double* __restrict__ partial,
                                                      KScalars* scs, const double* __restrict__ x,
                                                      const double* __restrict__ dinv) {
    KScalars* sc = scs + blockIdx.x;
    const Seg S = sg.seg[blockIdx.x];
    double tot[3];
    finish_sum<3>(partial, sg.nrows / BLOCK, S.blk0, S.nblk, tot);
    if (threadIdx.x != 0) return;
    double rr = tot[0], bb = tot[1];
    if (S.border) {
        const int nb_ = S.row0 + S.nv;
        const double rn = b[nb_] - tot[2];
        out[nb_] = rn;
        rr += rn * rn;
        bb += b[nb_] * b[nb_];
    }
    sc->red[0] = rr;
    sc->red[1] = bb;
    sc->red[2] = S.border ? x[S.row0 + S.nv] * dinv[S.row0 + S.nv] : 0.0;
}

// y = A x with the unscaled operator: A x = A~ (D x)
// mode 0: y = D^-1 x (d holds D^-1); mode 1: y = D x; mode 2: y = -x
__global__ __launch_bounds__(BLOCK) void k_scale(int n, const double* __restrict__ d,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ y, int mode) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (mode == 0)
        y[i] = x[i] * d[i];
    else if (mode == 1)
        y[i] = x[i] / d[i];
    else
        y[i] = -x[i];
}

// Resets the scalars of every system for a new run (a kernel rather than a
// memset, so the solve may sit in a captured graph): systems not taking part
// are marked done = 6.
__global__ void k_sc_reset(int nseg, KScalars* sc, const SegCtl* ctl) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nseg) return;
    KScalars z{};
    z.maxit = ctl[q].maxit;
    z.tol_rel = ctl[q].tol_rel;
    z.done = ctl[q].active ? 0 : 6;
    sc[q] = z;
}

// dst = src (the solver's device-to-device copies: a plain kernel streams
// them several times faster than the runtime's copy for these sizes)
__global__ __launch_bounds__(BLOCK) void k_copy_vec(int n, const double* __restrict__ src,
                                                     double* __restrict__ dst) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(BLOCK) void k_fill(double* __restrict__ x, double v, int n) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = v;
}

// y += a x
__global__ __launch_bounds__(BLOCK) void k_axpy(int n, double a, const double* __restrict__ x,
                                                double* __restrict__ y) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

// A~ = A D^-1 (column scaling) for the SELL slots; D^-1 of the border unknown is 1.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_scale_columns(SellView sv, const double* __restrict__ a,
                                                         const double* __restrict__ dinv,
                                                         T* __restrict__ out) {
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= sv.nv) return;
    typedef double d4 __attribute__((ext_vector_type(4)));
    typedef int i4 __attribute__((ext_vector_type(4)));
    typedef T t4 __attribute__((ext_vector_type(4)));
    const int base = sell_base(sv.sptr, r);
    const int len = sv.len[r];
    for (int j0 = 0; j0 < len; j0 += 4) {  // groups of four slots; past the length a = 0
        const int s4 = base + (j0 << 6);
        const i4 c = *reinterpret_cast<const i4*>(sv.col + s4);
        const d4 v = *reinterpret_cast<const d4*>(a + s4);
        const t4 o = {(T)(v.x * dinv[c.x]), (T)(v.y * dinv[c.y]), (T)(v.z * dinv[c.z]),
                      (T)(v.w * dinv[c.w])};
        *reinterpret_cast<t4*>(out + s4) = o;
    }
}

// out[s] = in[perm[s]] over the slots (values into the column-sorted order)
__global__ __launch_bounds__(BLOCK) void k_permute_slots(int n, const int* __restrict__ perm,
                                                         const float* __restrict__ in,
                                                         float* __restrict__ out) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}

// D^-1 of the operator's diagonal; rows without a diagonal (border unknowns,
// padding) get 1.
__global__ __launch_bounds__(BLOCK) void k_diag_inverse(int n, const double* __restrict__ diag,
                                                        double* __restrict__ dinv) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i];
    dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
}

inline int nblk(int n) { return (n + BLOCK - 1) / BLOCK; }

// ---- GMRES(m), right-preconditioned ----
// Per system: Arnoldi with classical Gram-Schmidt over FP64 basis vectors V
// (node rows and the border unknown), the Hessenberg column's Givens
// rotations and the residual estimate |g_j| on the device (one thread per
// system), the preconditioned directions Z_j = M v_j kept so that a cycle
// ends with x += Z y (no further V-cycle).  A system's kernels return at
// entry once it has stopped: converged, at its iteration cap, broken down,
// or with a full cycle waiting for its restart.  Partials: rows 0..GM_MAX-1
// of the work array hold the dot products, row GM_MAX the border row of A z.

__global__ void k_gm_reset(int nseg, GmState* gm, GmFlags* fl, const SegCtl* ctl) {
    const int q = blockIdx.x * 64 + threadIdx.x;
    if (q >= nseg) return;
    GmState* g = gm + q;
    g->done = ctl[q].active ? 0 : 6;
    g->stop = ctl[q].active ? 0 : 1;
    g->maxit = ctl[q].maxit;
    g->tol_rel = ctl[q].tol_rel;
    g->iter = 0;
    g->cycles = 0;
    g->j = 0;
    g->full = 0;
    g->apply = 0;
    g->napply = 0;
    fl[q] = GmFlags{g->done, 0, 0, 0};
}

// v0 = rhs - A x (unnormalised; the basis keeps it so, with sig[0] = 1/|v0|)
// for the systems starting (mode 0) or restarting (mode 1: a full cycle,
// still running), and the V-cycle's FP32 input (d0 v0, or v0 without the
// scaling) over the node rows; partial sums of v0.v0, rhs.rhs and the border
// row's D^-1 x.  x_zero: x is known to be zero (a defect-correction cycle's
// start), so A x is not formed.
template <typename TM, typename TV>
__global__ __launch_bounds__(BLOCK) void k_gm_residual(SellView sv, SegView sg,
                                                       const TM* __restrict__ a,
                                                       const double* __restrict__ dinv,
                                                       const double* __restrict__ rhs,
                                                       const double* __restrict__ x,
                                                       TV* __restrict__ v0,
                                                       const float* __restrict__ d0,
                                                       float* __restrict__ xs,
                                                       double* __restrict__ partial,
                                                       const GmState* gm, int mode, int x_zero) {
    const int nbk = sg.nrows / BLOCK;
    const int lb = xcd_block(nbk);
    if (lb < 0) return;
    const int q = block_seg(sg, lb);
    const GmState* g = gm + q;
    if (mode == 0 ? g->done != 0 : !(g->full && g->done == 0)) return;
    const Seg S = sg.seg[q];
    const int iend = S.row0 + S.nv;
    const int i = lb * BLOCK + threadIdx.x;
    const int s = wave_slice(lb);
    double vals[3] = {0.0, 0.0, 0.0};
    if (s < sv.nslices) {
        double ax = x_zero ? 0.0 : sell_dot<double>(sv, a, x, s, threadIdx.x & 63);
        if (i < iend) {
            if (S.border && !x_zero) ax = ax + x[iend];
            const double ri = rhs[i] - ax;
            const TV vi = (TV)ri;
            v0[i] = vi;
            if (xs) xs[i] = d0 ? d0[i] * (float)vi : (float)vi;
            vals[0] = ri * ri;
            vals[1] = rhs[i] * rhs[i];
            if (S.border && !x_zero) vals[2] = dinv[i] * x[i];
        }
    }
    block_store<3>(vals, partial, lb, nbk);
}

// (v0 holds the residual rounded to the basis type; its norm is the FP64 one)
template <typename TV>
__global__ __launch_bounds__(FIN) void k_gm_residual_finish(SegView sg,
                                                            const double* __restrict__ rhs,
                                                            TV* __restrict__ v0,
                                                            const double* __restrict__ partial,
                                                            GmState* gm, GmFlags* fl, int mode,
                                                            int rel_to_rhs) {
    GmState* g = gm + blockIdx.x;
    if (mode == 0 ? g->done != 0 : !(g->full && g->done == 0)) return;
    const Seg S = sg.seg[blockIdx.x];
    double tot[3];
    finish_sum<3>(partial, sg.nrows / BLOCK, S.blk0, S.nblk, tot);
    if (threadIdx.x != 0) return;
    double rr = tot[0], bb = tot[1];
    if (S.border) {
        const int ib = S.row0 + S.nv;
        const double rn = rhs[ib] - tot[2];
        v0[ib] = (TV)rn;
        rr += rn * rn;
        bb += rhs[ib] * rhs[ib];
    }
    if (mode == 0) {
        g->bb = bb;
        g->tol2 = g->tol_rel * g->tol_rel * (rel_to_rhs ? bb : rr);
        if (rel_to_rhs && bb == 0.0) g->done = 4;  // zero right-hand side: x = 0
        g->rr_cycle = rr;
    } else {
        g->cycles += 1;
        // restarted GMRES can stagnate (a non-normal operator, a weak
        // preconditioner): a cycle that reduced |r| by less than 10% stops
        // the system as stalled (done = 7; x keeps the cycle's result), for the
        // caller's fallback; a cycle that ended within the tolerance has
        // converged, however little it reduced |r| (a non-finite rr is caught
        // below)
        if (g->done == 0 && rr > g->tol2 && !(rr < 0.81 * g->rr_cycle) && rr <= 1.0e300)
            g->done = 7;
        g->rr_cycle = rr;
    }
    g->g[0] = sqrt(rr);
    g->rr = rr;
    g->inv_eta = 1.0 / g->g[0];
    g->sig[0] = g->inv_eta;
    g->j = 0;
    g->full = 0;
    g->apply = 0;
    if (g->done == 0) {
        if (!(rr <= 1.0e300))
            g->done = 5;
        else if (rr <= g->tol2)
            g->done = 1;
    }
    g->stop = g->done != 0;
    fl[blockIdx.x] = GmFlags{g->done, 0, g->iter, g->cycles};
}

// w = A z: node rows (plus the border unknown's column), partial sums of the
// border row sum D^-1 z; zg: the gathered copy of z (FP32 when z holds FP32
// values, the V-cycle's output).
template <typename TM, typename TX, typename TB>
__global__ __launch_bounds__(BLOCK) void k_gm_spmv(SellView sv, SegView sg,
                                                   const TM* __restrict__ a,
                                                   const double* __restrict__ dinv,
                                                   const TB* __restrict__ z,
                                                   const TX* __restrict__ zg,
                                                   double* __restrict__ w,
                                                   double* __restrict__ partial,
                                                   const GmState* gm) {
    const int nbk = sg.nrows / BLOCK;
    const int lb = xcd_block(nbk);
    if (lb < 0) return;
    const int q = block_seg(sg, lb);
    if (gm[q].stop) return;
    const Seg S = sg.seg[q];
    const int iend = S.row0 + S.nv;
    const int i = lb * BLOCK + threadIdx.x;
    const int s = wave_slice(lb);
    double vals[1] = {0.0};
    if (s < sv.nslices) {
        double y = sell_dot<double>(sv, a, zg, s, threadIdx.x & 63);
        if (i < iend) {
            if (S.border) y = y + (double)z[iend];
            w[i] = y;
            if (S.border) vals[0] = dinv[i] * (double)zg[i];
        }
    }
    block_store<1>(vals, partial + (size_t)GM_MAX * nbk, lb, nbk);
}

// partial sums of v_k . w over the node rows, k < nk (the border row is added
// by the finishing kernel once w's border entry is known).  The block's 256
// rows are shared by its four wavefronts; wavefront u forms the dots of
// v_u, v_{u+4}, ... (four rows per lane, then one shuffle tree per vector).
template <typename TV>
__global__ __launch_bounds__(BLOCK) void k_gm_dots(SegView sg, const double* __restrict__ w,
                                                   const TV* __restrict__ V, size_t ld,
                                                   int nk, double* __restrict__ partial,
                                                   const GmState* gm) {
    static_assert(BLOCK == 256, "four rows per lane");
    const int nbk = sg.nrows / BLOCK;
    const int lb = blockIdx.x;
    const int q = block_seg(sg, lb);
    if (gm[q].stop) return;
    const Seg S = sg.seg[q];
    const int iend = S.row0 + S.nv;
    const int lane = threadIdx.x & 63, u = threadIdx.x >> 6;
    const int r0 = lb * BLOCK + lane;
    double wr[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) wr[t] = (r0 + 64 * t < iend) ? w[r0 + 64 * t] : 0.0;
    for (int k = u; k < nk; k += 4) {
        const TV* vk = V + k * ld + r0;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (r0 + 64 * t < iend) acc = fma((double)vk[64 * t], wr[t], acc);
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) partial[(size_t)k * nbk + lb] = acc;
    }
}

// Hessenberg column: h_k = sum of the partials of V'_k . w plus the border
// term, scaled to the normalised vectors (w = A M V'_j is sig[j] times the
// product of v_j): h_k = sig[k] sig[j] (V'_k . w), and V'_k's coefficient in the
// Gram-Schmidt update hc[k] = h_k sig[k]; w's border entry from its partial
// sums.  One block per (system, k): each sums its row of block partials and
// the border row (every 256th partial per thread, four in flight, then a
// fixed reduction tree).
constexpr int GMF = 256;  // threads of the Hessenberg-column kernel
__device__ __forceinline__ double gm_row_sum(const double* __restrict__ p, int cnt) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = threadIdx.x;
    for (; b + 3 * GMF < cnt; b += 4 * GMF) {
        a0 += p[b];
        a1 += p[b + GMF];
        a2 += p[b + 2 * GMF];
        a3 += p[b + 3 * GMF];
    }
    for (; b < cnt; b += GMF) a0 += p[b];
    double acc = (a0 + a1) + (a2 + a3);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    return acc;  // per wavefront
}

template <typename TV>
__global__ __launch_bounds__(GMF) void k_gm_finish_h(SegView sg, double* __restrict__ w,
                                                     const TV* __restrict__ V, size_t ld,
                                                     int nk, const double* __restrict__ partial,
                                                     GmState* gm) {
    const int q = blockIdx.x / GM_MAX, k = blockIdx.x % GM_MAX;
    GmState* g = gm + q;
    if (k >= nk || g->stop) return;
    const Seg S = sg.seg[q];
    const int nb = sg.nrows / BLOCK;
    __shared__ double sh[2][GMF / 64];
    const double dk = gm_row_sum(partial + (size_t)k * nb + S.blk0, S.nblk);
    const double db = S.border ? gm_row_sum(partial + (size_t)GM_MAX * nb + S.blk0, S.nblk) : 0.0;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wid] = dk;
        sh[1][wid] = db;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tk = sh[0][0], wb = sh[1][0];
    for (int t = 1; t < GMF / 64; ++t) {
        tk += sh[0][t];
        wb += sh[1][t];
    }
    if (S.border) {
        const int ib = S.row0 + S.nv;
        if (k == 0) w[ib] = wb;
        tk += (double)V[k * ld + ib] * wb;
    }
    const double h = g->sig[k] * g->sig[nk - 1] * tk;
    g->h[k] = h;
    g->hc[k] = h * g->sig[k];
}

// The next basis vector, unnormalised: V'_nk = sig[nk-1] w - sum_k hc[k] V'_k
// over the node rows and the border unknown (rounded to the basis type), the
// V-cycle's FP32 input (d0 V'_nk, or V'_nk without the scaling) over the node
// rows, and partial sums of V'_nk . V'_nk (its 1/norm becomes sig[nk]).
template <typename TV>
__global__ __launch_bounds__(BLOCK) void k_gm_orth(SegView sg, const double* __restrict__ w,
                                                   TV* __restrict__ V, size_t ld, int nk,
                                                   const float* __restrict__ d0,
                                                   float* __restrict__ xs,
                                                   double* __restrict__ partial,
                                                   const GmState* gm) {
    const int nbk = sg.nrows / BLOCK;
    const int lb = blockIdx.x;
    const int q = block_seg(sg, lb);
    const GmState* g = gm + q;
    if (g->stop) return;
    const Seg S = sg.seg[q];
    const int i = lb * BLOCK + threadIdx.x;
    double vals[1] = {0.0};
    if (i < S.row0 + S.nv + S.border) {
        double wi = g->sig[nk - 1] * w[i];
        for (int k = 0; k < nk; ++k) wi = fma(-g->hc[k], (double)V[k * ld + i], wi);
        const TV v = (TV)wi;
        V[(size_t)nk * ld + i] = v;
        if (xs && i < S.row0 + S.nv) xs[i] = d0 ? d0[i] * (float)v : (float)v;
        vals[0] = (double)v * (double)v;
    }
    block_store<1>(vals, partial, lb, nbk);
}

// Givens rotation (c, s) with c a + s b = r, -s a + c b = 0
__device__ __forceinline__ void gm_givens(double a, double b, double& c, double& s, double& r) {
    if (b == 0.0) {
        c = 1.0;
        s = 0.0;
        r = a;
    } else {
        r = hypot(a, b);
        c = a / r;
        s = b / r;
    }
}

// h_{j+1,j} = |w|; the column's rotations, the new rotation and the residual
// estimate; then the stopping tests.
__global__ __launch_bounds__(FIN) void k_gm_finish_norm(SegView sg,
                                                        const double* __restrict__ partial,
                                                        GmState* gm, GmFlags* fl, int m) {
    GmState* g = gm + blockIdx.x;
    if (g->stop) return;
    const Seg S = sg.seg[blockIdx.x];
    double tot[1];
    finish_sum<1>(partial, sg.nrows / BLOCK, S.blk0, S.nblk, tot);
    // the column and the previous rotations into LDS in one parallel load;
    // the rotations themselves are sequential (thread 0)
    __shared__ double h[GM_MAX + 1], cs[GM_MAX], sn[GM_MAX];
    const int j = g->j;
    if ((int)threadIdx.x <= j) h[threadIdx.x] = g->h[threadIdx.x];
    if ((int)threadIdx.x < j) {
        cs[threadIdx.x] = g->cs[threadIdx.x];
        sn[threadIdx.x] = g->sn[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < j; ++i) {
            const double a = h[i], b = h[i + 1];
            h[i] = cs[i] * a + sn[i] * b;
            h[i + 1] = -sn[i] * a + cs[i] * b;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < j) g->R[j * GM_MAX + threadIdx.x] = h[threadIdx.x];
    if (threadIdx.x != 0) return;
    const double eta = sqrt(tot[0]);
    double c, s, r;
    gm_givens(h[j], eta, c, s, r);
    g->cs[j] = c;
    g->sn[j] = s;
    g->R[j * GM_MAX + j] = r;
    g->g[j + 1] = -s * g->g[j];
    g->g[j] = c * g->g[j];
    g->j = j + 1;
    g->iter += 1;
    const double res2 = g->g[j + 1] * g->g[j + 1];
    g->rr = res2;
    g->inv_eta = 1.0 / eta;
    g->sig[j + 1] = g->inv_eta;
    if (!(res2 <= 1.0e300) || !(fabs(r) <= 1.0e300))
        g->done = 5;
    else if (res2 <= g->tol2)
        g->done = 1;
    else if (r == 0.0)
        g->done = 5;  // singular Hessenberg column
    else if (g->iter >= g->maxit)
        g->done = 2;
    else if (g->j >= m)
        g->full = 1;
    g->stop = (g->done != 0 || g->full) ? 1 : 0;
    fl[blockIdx.x] = GmFlags{g->done, g->full, g->iter, g->cycles};
}

// end of a cycle: y = R^-1 g over the cycle's columns (back substitution)
__global__ __launch_bounds__(64) void k_gm_solve(GmState* gm, GmFlags* fl) {
    // one block per system: the cycle's triangle and g are staged in LDS,
    // then one thread substitutes (the same operations in the same order)
    __shared__ double sR[GM_MAX * GM_MAX], sg[GM_MAX + 1], sy[GM_MAX];
    const int q = blockIdx.x;
    GmState* g = gm + q;
    const int j = g->j;
    const bool run = j > 0 && g->done < 4;
    if (run) {
        for (int t = threadIdx.x; t < j * GM_MAX; t += 64) sR[t] = g->R[t];
        if ((int)threadIdx.x < j) sg[threadIdx.x] = g->g[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    g->apply = 0;
    g->napply = 0;
    g->j = 0;  // the cycle's columns are consumed here, once
    if (!run) return;
    bool finite = true;
    for (int i = j - 1; i >= 0; --i) {
        double t = sg[i];
        for (int k = i + 1; k < j; ++k) t -= sR[k * GM_MAX + i] * sy[k];
        sy[i] = t / sR[i * GM_MAX + i];
        finite = finite && fabs(sy[i]) <= 1.0e300;
    }
    for (int i = 0; i < j; ++i) g->y[i] = sy[i];
    if (!finite) {
        g->done = 5;
        g->stop = 1;
        fl[q].done = 5;
        return;
    }
    g->apply = 1;
    g->napply = j;
}

// x += sum_k y_k z_k over the node rows and the border unknown, with
// z_k = sig[k] Z'_k, Z'_k = M V'_k (z_k's border entry is v_k's: the
// preconditioner is the identity on the border unknown)
template <typename TZ, typename TV>
__global__ __launch_bounds__(BLOCK) void k_gm_update_x(SegView sg, double* __restrict__ x,
                                                       const TZ* __restrict__ Z,
                                                       const TV* __restrict__ V, size_t ld,
                                                       const GmState* gm) {
    const int q = block_seg(sg, blockIdx.x);
    const GmState* g = gm + q;
    if (!g->apply) return;
    const Seg S = sg.seg[q];
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const int iend = S.row0 + S.nv;
    if (i >= iend + S.border) return;
    double xi = x[i];
    if (i < iend)
        for (int k = 0; k < g->napply; ++k) xi = fma(g->y[k] * g->sig[k], (double)Z[k * ld + i], xi);
    else
        for (int k = 0; k < g->napply; ++k) xi = fma(g->y[k] * g->sig[k], (double)V[k * ld + i], xi);
    x[i] = xi;
}

}  // namespace

void launch_fill(hipStream_t s, double* x, double v, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(BLOCK), 0, s, x, v, n);
    MP_HIP(hipGetLastError());
}

void launch_copy(hipStream_t s, double* dst, const double* src, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_vec, dim3(nblk(n)), dim3(BLOCK), 0, s, n, src, dst);
    MP_HIP(hipGetLastError());
}

void KernelTimer::begin(hipStream_t s) {
    if (!enabled) return;
    if (used + 2 > ev.size()) {
        size_t old = ev.size();
        ev.resize(std::max<size_t>(64, old * 2));
        for (size_t i = old; i < ev.size(); ++i) MP_HIP(hipEventCreate(&ev[i]));
    }
    MP_HIP(hipEventRecord(ev[used], s));
}
void KernelTimer::end(hipStream_t s) {
    if (!enabled) return;
    MP_HIP(hipEventRecord(ev[used + 1], s));
    used += 2;
}
void KernelTimer::harvest() {
    for (size_t i = 0; i + 1 < used; i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess && ms > 0.0f)
            samples.push_back(ms);
    }
    used = 0;
}
void KernelTimer::summary(long* working, double* mean_ms, long* all) const {
    *working = 0;
    *mean_ms = 0.0;
    *all = (long)samples.size();
    if (samples.empty()) return;
    // (a quarter of the longest launch: launches that returned at once are
    // a few microseconds, working ones tens; the median fails to separate
    // them when most launches of a short solve return at once)
    const float cut = 0.25f * *std::max_element(samples.begin(), samples.end());
    double tot = 0.0;
    for (float ms : samples)
        if (ms >= cut) {
            tot += ms;
            ++*working;
        }
    *mean_ms = tot / (double)*working;
}
void KernelTimer::release() {
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    used = 0;
}

PpeSolver::~PpeSolver() {
    release_graphs();
    if (host_sc) (void)hipHostFree(host_sc);
    if (host_ctl) (void)hipHostFree(host_ctl);
    if (host_gmf) (void)hipHostFree(host_gmf);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    spmv_timer.release();
    smooth_timer.release();
}

void PpeSolver::setup(hipStream_t st, const SellView& view, const std::vector<Seg>& systems,
                      int dim_, const double* gx, const double* gy, const double* gz,
                      const double* lap, const double* normal_soa, const uint8_t* bnd_dev,
                      const uint8_t* dir_dev, double unsteady, double conv, double diff,
                      int precond_) {
    release_graphs();
    gm_cols = 0;  // (the GMRES work arrays are sized for the rows set here)
    sv = view;
    nrows = view.nv;
    dim = dim_;
    segs = systems;
    nseg = (int)segs.size();
    if (nrows % BLOCK != 0 || nseg < 1) throw std::runtime_error("PpeSolver: bad row layout");
    nv = 0;
    n_sys = 0;
    bordered = 0;
    std::vector<int> hb(nrows / BLOCK, 0);
    for (int q = 0; q < nseg; ++q) {
        const Seg& S = segs[q];
        if (S.blk0 * BLOCK != S.row0 || S.row0 + S.nv + S.border > (S.blk0 + S.nblk) * BLOCK)
            throw std::runtime_error("PpeSolver: system not aligned to its row blocks");
        for (int b = 0; b < S.nblk; ++b) hb[S.blk0 + b] = q;
        nv += S.nv;
        n_sys += S.nv + S.border;
        bordered |= S.border;
    }
    seg_d.alloc(nseg);
    bseg_d.alloc(hb.size());
    MP_HIP(hipMemcpyAsync(seg_d.p, segs.data(), sizeof(Seg) * nseg, hipMemcpyHostToDevice, st));
    MP_HIP(hipMemcpyAsync(bseg_d.p, hb.data(), sizeof(int) * hb.size(), hipMemcpyHostToDevice, st));
    sg = SegView{nseg, nrows, seg_d.p, bseg_d.p};
    const size_t vlen = (size_t)nrows + 64;
    DevBuf<double> raw, diag;
    const int total_slots = [&]() {
        int t = 0;
        MP_HIP(hipMemcpy(&t, view.sptr + view.nslices, sizeof(int), hipMemcpyDeviceToHost));
        return t;
    }();
    raw.alloc(total_slots);
    diag.alloc(vlen);
    aval.alloc(total_slots);
    aval32.alloc(total_slots);
    dinv.alloc(vlen);
    MP_HIP(hipMemsetAsync(raw.p, 0, sizeof(double) * total_slots, st));
    MP_HIP(hipMemsetAsync(aval.p, 0, sizeof(double) * total_slots, st));
    MP_HIP(hipMemsetAsync(aval32.p, 0, sizeof(float) * total_slots, st));
    MP_HIP(hipMemsetAsync(diag.p, 0, sizeof(double) * vlen, st));
    launch_build_operator(st, sv, dim, gx, gy, gz, lap, normal_soa, nrows, bnd_dev, dir_dev,
                          unsteady, conv, diff, raw.p, diag.p);
    hipLaunchKernelGGL(k_diag_inverse, dim3(nblk((int)vlen)), dim3(BLOCK), 0, st, (int)vlen,
                       diag.p, dinv.p);
    hipLaunchKernelGGL(k_scale_columns<double>, dim3(nblk(nrows)), dim3(BLOCK), 0, st, sv, raw.p,
                       dinv.p, aval.p);
    hipLaunchKernelGGL(k_scale_columns<float>, dim3(nblk(nrows)), dim3(BLOCK), 0, st, sv, raw.p,
                       dinv.p, aval32.p);
    MP_HIP(hipGetLastError());
    precond = precond_;
    amg.reset();
    diag_inv = dinv.p;
    scaled32 = aval32.p;
    if (precond == MP_PRECOND_AMG) {
        // keep A D^-1 (FP32) and D^-1 for the learned guess, then switch the
        // Krylov operator to the unscaled A with a scaling vector of ones
        jac_dinv.alloc(vlen);
        jac_a32.alloc(total_slots);
        MP_HIP(hipMemcpyAsync(jac_dinv.p, dinv.p, sizeof(double) * vlen, hipMemcpyDeviceToDevice, st));
        MP_HIP(hipMemcpyAsync(jac_a32.p, aval32.p, sizeof(float) * total_slots,
                              hipMemcpyDeviceToDevice, st));
        launch_fill(st, dinv.p, 1.0, (int)vlen);
        hipLaunchKernelGGL(k_scale_columns<double>, dim3(nblk(nrows)), dim3(BLOCK), 0, st, sv,
                           raw.p, dinv.p, aval.p);
        hipLaunchKernelGGL(k_scale_columns<float>, dim3(nblk(nrows)), dim3(BLOCK), 0, st, sv,
                           raw.p, dinv.p, aval32.p);
        MP_HIP(hipGetLastError());
        diag_inv = jac_dinv.p;
        scaled32 = jac_a32.p;
    }
    // the FP32 operator of the Krylov products (and of the multigrid's finest
    // level) in the column-sorted slot order when there is one
    svk = sv;
    a32k = aval32.p;
    if (slot_perm) {
        aval32s.alloc(total_slots);
        hipLaunchKernelGGL(k_permute_slots, dim3(nblk(total_slots)), dim3(BLOCK), 0, st,
                           total_slots, slot_perm, aval32.p, aval32s.p);
        MP_HIP(hipGetLastError());
        svk = svs;
        svk.nv = nrows;
        a32k = aval32s.p;
    }
    if (precond == MP_PRECOND_AMG) {
        // one problem: the hierarchy is built on its node block (rows 0..nv);
        // a batch: one hierarchy per problem over all device rows
        SellView s0 = sv, s0k = svk;
        s0.nv = s0k.nv = nseg == 1 ? segs[0].nv : nrows;
        s0.nslices = s0k.nslices = (s0.nv + SLICE - 1) / SLICE;
        std::vector<std::pair<int, int>> blocks;
        std::vector<int> border_rows, border_flag;
        for (const Seg& S : segs) {
            blocks.emplace_back(S.row0, S.nv);
            border_flag.push_back(S.border);
            if (S.border) border_rows.push_back(S.row0 + S.nv);
        }
        amg.reset(new AmgDevice());
        amg->setup(st, s0, aval.p, s0k, a32k, blocks, border_flag, mp_amg::Options{});
        amg->post_timer = &smooth_timer;
        if (nseg > 1) amg->set_border_rows(st, border_rows);
        for (DevBuf<double>* b : {&yv, &zv}) {
            b->alloc(vlen);
            MP_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * vlen, st));
        }
        for (DevBuf<float>* b : {&y32, &z32}) {
            b->alloc(vlen);
            MP_HIP(hipMemsetAsync(b->p, 0, sizeof(float) * vlen, st));
        }
    }
    for (DevBuf<double>* b : {&xt, &r, &r0, &p, &v, &s, &t, &tmp, &e64, &rhs64}) {
        b->alloc(vlen);
        MP_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * vlen, st));
    }
    partial.alloc((size_t)(nrows / BLOCK) * 4);
    sc.alloc(nseg);
    MP_HIP(hipMemsetAsync(sc.p, 0, sizeof(KScalars) * nseg, st));
    ctl_d.alloc(nseg);
    if (host_sc) (void)hipHostFree(host_sc);
    if (host_ctl) (void)hipHostFree(host_ctl);
    host_sc = nullptr;
    host_ctl = nullptr;
    MP_HIP(hipHostMalloc((void**)&host_sc, 2 * sizeof(KScalars) * nseg, 0));
    MP_HIP(hipHostMalloc((void**)&host_ctl, sizeof(SegCtl) * nseg, 0));
    for (auto& e : ev)
        if (!e) MP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // nnz of the operator (real entries): sum of row lengths
    std::vector<int> hl(nrows);
    MP_HIP(hipMemcpyAsync(hl.data(), sv.len, sizeof(int) * nrows, hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
    nnz = 0;
    for (int i = 0; i < nrows; ++i) nnz += hl[i];
    nnz16 = 0;
    nslices16 = 0;
    if (svk.col16) {
        std::vector<int> cb((size_t)svk.nslices * CW_COUNT);
        MP_HIP(hipMemcpy(cb.data(), svk.cbase, sizeof(int) * cb.size(), hipMemcpyDeviceToHost));
        for (int sl = 0; sl < svk.nslices; ++sl)
            if (cb[(size_t)sl * CW_COUNT] >= 0) {
                ++nslices16;
                for (int l = 0; l < SLICE; ++l) nnz16 += hl[(size_t)sl * SLICE + l];
            }
    }
    seg_stats.assign(nseg, SolveStats{});
}

double PpeSolver::smooth_bytes() const {
    // per entry: BF16 value + column (2 B in slices with 16-bit columns, plus
    // their window bases, else 4 B); per row: the gathered FP32 input, b read,
    // D^-1 (FP32), the output, length.  GMRES: b is the FP32 basis vector and
    // the output the FP32 direction; BiCGSTAB: b FP64, the output in FP64 and
    // FP32
    if (!amg) return 0.0;
    const double n16 = (double)nnz16;
    const double per_row = krylov == 1 ? 4.0 + 4.0 + 4.0 + 4.0 + 4.0
                                       : 4.0 + 8.0 + 4.0 + 8.0 + 4.0 + 4.0;
    return (double)nnz * 2.0 + 2.0 * n16 + 4.0 * ((double)nnz - n16) +
           (double)nslices16 * 4.0 * CW_COUNT + (double)nv * per_row;
}

double PpeSolver::spmv_bytes(int fp32) const {
    // per entry: value (8 B, or 4 B for the FP32 operator) + column (4 B; the
    // FP32 operator: 2 B in slices with 16-bit columns, plus their window
    // bases); per row: gathered x once (FP32 with the V-cycle, else FP64),
    // the product written, r0 read (BiCGSTAB's fused dot; FP64), length,
    // D^-1 for the border row
    const double vb = fp32 ? 4.0 : 8.0;
    const double n16 = fp32 ? (double)nnz16 : 0.0;
    // the V-cycle's output, and an FP32 GMRES basis vector, are gathered in FP32
    const double xb = (amg || (krylov == 1 && fp32)) ? 4.0 : 8.0;
    const double vec = (krylov == 1 ? 1.0 : 2.0) + (bordered ? 1.0 : 0.0);
    return (double)nnz * vb + 2.0 * n16 + 4.0 * ((double)nnz - n16) +
           (fp32 ? (double)nslices16 * 4.0 * CW_COUNT : 0.0) +
           (double)nv * (xb + 8.0 * vec + 4.0);
}

// y = M^-1 p with the multigrid V-cycle; without it the direction itself.
// The preceding update kernel has written the V-cycle's scaled FP32 input.
const double* PpeSolver::precondition(hipStream_t st, const double* in, double* out,
                                      float* out32) {
    if (!amg) return in;
    // one problem: its kernels stop once it has converged; a batch runs the
    // V-cycle while any problem still iterates (the others ignore it)
    amg->apply(st, in, out, nseg == 1 ? segs[0].border : 1, nseg == 1 ? &sc.p->done : nullptr,
               amg->prescaled_input() != nullptr, out32);
    return out;
}

// The chunk of iterations between two convergence polls, launched on the
// stream (captured into a graph by run_chunk).
template <typename TM, typename T>
void PpeSolver::launch_chunk(hipStream_t st, int chunk, const TM* a, const T* di, T* x, T* rr,
                             T* rr0, T* pp, T* vv, T* ss, T* tt) {
    const int nbk = nrows / BLOCK;
    float* xs = amg ? amg->prescaled_input() : nullptr;
    const float* d0 = amg ? amg->level0_dinv() : nullptr;
    for (int c = 0; c < chunk; ++c) {
        hipLaunchKernelGGL(k_update_p<T>, dim3(nbk), dim3(BLOCK), 0, st, sg, rr, pp, vv, sc.p, d0,
                           xs);
        // with the V-cycle, the products gather its FP32 output (the same values)
        const T* yin = precondition(st, pp, yv.p, y32.p);
        spmv_timer.begin(st);
        if (amg)
            hipLaunchKernelGGL((k_spmv_alpha<TM, T, float>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0,
                               st, view_of<TM>(), sg, a, di, yin, (const float*)y32.p, vv, rr0,
                               partial.p, sc.p);
        else
            hipLaunchKernelGGL((k_spmv_alpha<TM, T, T>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st,
                               view_of<TM>(), sg, a, di, yin, yin, vv, rr0, partial.p, sc.p);
        spmv_timer.end(st);
        hipLaunchKernelGGL(k_finish_alpha<T>, dim3(nseg), dim3(FIN), 0, st, sg, vv, rr0, partial.p,
                           sc.p);
        hipLaunchKernelGGL(k_update_s<T>, dim3(nbk), dim3(BLOCK), 0, st, sg, rr, vv, ss, sc.p, d0,
                           xs);
        const T* zin = precondition(st, ss, zv.p, z32.p);
        spmv_timer.begin(st);
        if (amg)
            hipLaunchKernelGGL((k_spmv_omega<TM, T, float>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0,
                               st, view_of<TM>(), sg, a, di, zin, (const float*)z32.p, ss, tt,
                               partial.p, sc.p);
        else
            hipLaunchKernelGGL((k_spmv_omega<TM, T, T>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st,
                               view_of<TM>(), sg, a, di, zin, zin, ss, tt, partial.p, sc.p);
        spmv_timer.end(st);
        hipLaunchKernelGGL(k_finish_omega<T>, dim3(nseg), dim3(FIN), 0, st, sg, tt, ss, partial.p,
                           sc.p);
        hipLaunchKernelGGL(k_update_xr<T>, dim3(nbk), dim3(BLOCK), 0, st, sg, x, rr, yin, zin, ss,
                           tt, rr0, partial.p, sc.p);
        hipLaunchKernelGGL(k_finish_xr, dim3(nseg), dim3(FIN), 0, st, sg, partial.p, sc.p);
    }
    MP_HIP(hipGetLastError());
}

// Replays the chunk from a cached graph (captured on first use for this set of
// operands); per-launch event timing needs the plain launches instead.
template <typename TM, typename T>
void PpeSolver::run_chunk(hipStream_t st, int chunk, const TM* a, const T* di, T* x, T* rr,
                          T* rr0, T* pp, T* vv, T* ss, T* tt) {
    if (!use_graphs || spmv_timer.enabled || smooth_timer.enabled) {  // (events stay out of graphs)
        launch_chunk<TM, T>(st, chunk, a, di, x, rr, rr0, pp, vv, ss, tt);
        return;
    }
    const uintptr_t key[ChunkGraph::KEY] = {
        1, sizeof(TM) * 16 + sizeof(T), (uintptr_t)chunk, (uintptr_t)a, (uintptr_t)di,
        (uintptr_t)x, (uintptr_t)rr, (uintptr_t)rr0, (uintptr_t)pp, (uintptr_t)vv,
        (uintptr_t)ss, (uintptr_t)tt};
    replay_or_capture(st, key,
                      [&]() { launch_chunk<TM, T>(st, chunk, a, di, x, rr, rr0, pp, vv, ss, tt); });
}

template <typename F>
void PpeSolver::replay_or_capture(hipStream_t st, const uintptr_t (&key)[ChunkGraph::KEY],
                                  F&& launch) {
    for (ChunkGraph& g : graphs)
        if (std::memcmp(g.key, key, sizeof(key)) == 0) {
            MP_HIP(hipGraphLaunch(g.exec, st));
            return;
        }
    hipGraph_t graph = nullptr;
    MP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        launch();
    } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    MP_HIP(hipStreamEndCapture(st, &graph));
    ChunkGraph g{};
    std::memcpy(g.key, key, sizeof(key));
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MP_HIP(e);
    graphs.push_back(g);
    MP_HIP(hipGraphLaunch(g.exec, st));
}

template <typename TM, typename TV>
void PpeSolver::run_gmres_chunk(hipStream_t st, int j, int c, int m, const TM* a,
                                const double* di) {
    if (!use_graphs || spmv_timer.enabled || smooth_timer.enabled) {  // (events stay out of graphs)
        for (int k = j; k < j + c; ++k) gmres_iteration<TM, TV>(st, k, m, a, di);
        return;
    }
    const uintptr_t key[ChunkGraph::KEY] = {
        2, sizeof(TM) * 16 + sizeof(TV), (uintptr_t)j, (uintptr_t)c, (uintptr_t)m,
        (uintptr_t)a, (uintptr_t)di, (uintptr_t)gm_v.p, (uintptr_t)gm_d.p,
        (uintptr_t)gm_partial.p, (uintptr_t)gm_w.p, (uintptr_t)amg.get()};
    replay_or_capture(st, key, [&]() {
        for (int k = j; k < j + c; ++k) gmres_iteration<TM, TV>(st, k, m, a, di);
    });
}

void PpeSolver::release_graphs() {
    for (ChunkGraph& g : graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

// One BiCGSTAB run for the systems with ctl[q].active (controls already in
// host_ctl): fills done[q], iters[q], restarts[q] of the taking part systems.
template <typename TM, typename T>
void PpeSolver::run_bicgstab(hipStream_t st, const TM* a, const T* di, const T* rhs, T* x, T* rr,
                             T* rr0, T* pp, T* vv, T* ss, T* tt, int rel_to_rhs,
                             std::vector<int>& done, std::vector<int>& iters,
                             std::vector<int>& restarts) {
    const int nbk = nrows / BLOCK;
    MP_HIP(hipMemcpyAsync(ctl_d.p, host_ctl, sizeof(SegCtl) * nseg, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_sc_reset, dim3((nseg + 63) / 64), dim3(64), 0, st, nseg, sc.p, ctl_d.p);
    hipLaunchKernelGGL((k_residual_init<TM, T>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st,
                       view_of<TM>(), sg,
                       a, di, rhs, x, rr, rr0, pp, vv, partial.p, sc.p, 0);
    hipLaunchKernelGGL(k_finish_init<T>, dim3(nseg), dim3(FIN), 0, st, sg, rhs, rr, rr0, pp, vv,
                       partial.p, sc.p, 0, rel_to_rhs);
    MP_HIP(hipGetLastError());
    const int chunk = std::max(1, check_every);
    int slot = 0;
    bool have_prev = false;
    KScalars* hs[2] = {host_sc, host_sc + nseg};
    while (true) {
        run_chunk<TM, T>(st, chunk, a, di, x, rr, rr0, pp, vv, ss, tt);
        MP_HIP(hipMemcpyAsync(hs[slot], sc.p, sizeof(KScalars) * nseg, hipMemcpyDeviceToHost, st));
        MP_HIP(hipEventRecord(ev[slot], st));
        if (have_prev) {
            const int ps = slot ^ 1;
            MP_HIP(hipEventSynchronize(ev[ps]));
            bool restart = false, running = false;
            for (int q = 0; q < nseg; ++q) {
                restart |= hs[ps][q].done == 3;
                running |= hs[ps][q].done == 0;
            }
            if (restart) {
                // breakdown of r0.r in some system: once the in-flight chunk has
                // run, restart those systems with r0 = r = rhs - A x, as Eigen's
                // BiCGSTAB does (the kernels act on the systems still at done == 3)
                MP_HIP(hipEventSynchronize(ev[slot]));
                hipLaunchKernelGGL((k_residual_init<TM, T>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0,
                                   st, view_of<TM>(), sg, a, di, rhs, x, rr, rr0, pp, vv,
                                   partial.p, sc.p, 1);
                hipLaunchKernelGGL(k_finish_init<T>, dim3(nseg), dim3(FIN), 0, st, sg, rhs, rr,
                                   rr0, pp, vv, partial.p, sc.p, 1, rel_to_rhs);
                MP_HIP(hipGetLastError());
                have_prev = false;
                slot ^= 1;
                continue;
            }
            if (!running) {
                MP_HIP(hipEventSynchronize(ev[slot]));
                break;
            }
        }
        have_prev = true;
        slot ^= 1;
    }
    for (int q = 0; q < nseg; ++q) {
        const KScalars& h = hs[slot][q];
        if (!host_ctl[q].active) continue;
        done[q] = h.done;
        iters[q] = h.iter;
        restarts[q] = h.restarts;
    }
}

// Per-system FP64 residual |b - A x~|^2 and |b|^2 of xt into tmp; waits.
void PpeSolver::residual64(hipStream_t st, const double* b_dev, std::vector<double>& rr,
                           std::vector<double>& bb) {
    const int nbk = nrows / BLOCK;
    hipLaunchKernelGGL(k_residual64, dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st, sv, sg, aval.p,
                       dinv.p, b_dev, xt.p, tmp.p, partial.p);
    hipLaunchKernelGGL(k_finish_res64, dim3(nseg), dim3(FIN), 0, st, sg, b_dev, tmp.p, partial.p,
                       sc.p, xt.p, dinv.p);
    MP_HIP(hipGetLastError());
    MP_HIP(hipMemcpyAsync(host_sc, sc.p, sizeof(KScalars) * nseg, hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
    rr.resize(nseg);
    bb.resize(nseg);
    for (int q = 0; q < nseg; ++q) {
        rr[q] = host_sc[q].red[0];
        bb[q] = host_sc[q].red[1];
    }
}

double PpeSolver::last_residual_norm2(hipStream_t st) {
    MP_HIP(hipMemcpyAsync(host_sc, sc.p, sizeof(KScalars), hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
    return host_sc[0].red[0];
}

double* PpeSolver::residual_of(hipStream_t st, const double* b_dev, const double* x_dev) {
    const int nbk = nrows / BLOCK;
    hipLaunchKernelGGL(k_scale, dim3(nbk), dim3(BLOCK), 0, st, nrows, dinv.p, x_dev, xt.p, 1);
    hipLaunchKernelGGL(k_residual64, dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st, sv, sg, aval.p,
                       dinv.p, b_dev, xt.p, tmp.p, partial.p);
    hipLaunchKernelGGL(k_finish_res64, dim3(nseg), dim3(FIN), 0, st, sg, b_dev, tmp.p, partial.p,
                       sc.p, xt.p, dinv.p);
    MP_HIP(hipGetLastError());
    return tmp.p;
}

void PpeSolver::gmres_alloc(hipStream_t st, int m) {
    if (gm_cols == m && gm_d.p) return;
    release_graphs();  // (captured chunks hold the buffers replaced here)
    const size_t ld = (size_t)nrows;
    gm_v.alloc(ld * (m + 1));
    MP_HIP(hipMemsetAsync(gm_v.p, 0, sizeof(double) * ld * (m + 1), st));
    if (amg) {
        gm_z.alloc(ld * m);
        MP_HIP(hipMemsetAsync(gm_z.p, 0, sizeof(float) * ld * m, st));
    }
    gm_w.alloc(ld);
    MP_HIP(hipMemsetAsync(gm_w.p, 0, sizeof(double) * ld, st));
    gm_partial.alloc((size_t)(GM_MAX + 1) * (nrows / BLOCK));
    gm_d.alloc(nseg);
    MP_HIP(hipMemsetAsync(gm_d.p, 0, sizeof(GmState) * nseg, st));
    gmf_d.alloc(nseg);
    if (host_gmf) (void)hipHostFree(host_gmf);
    host_gmf = nullptr;
    MP_HIP(hipHostMalloc((void**)&host_gmf, 2 * sizeof(GmFlags) * nseg, 0));
    gm_cols = m;
}

// Iteration k of a GMRES cycle: z_k = M v_k, w = A z_k, Gram-Schmidt against
// v_0..v_k, the Hessenberg column's rotations, v_{k+1} = w / |w|.
template <typename TM, typename TV>
void PpeSolver::gmres_iteration(hipStream_t st, int k, int m, const TM* a, const double* di) {
    const int nbk = nrows / BLOCK;
    const size_t ld = (size_t)nrows;
    TV* V = reinterpret_cast<TV*>(gm_v.p);
    TV* vk = V + ld * k;
    if (amg) {
        // one system: its V-cycle stops with it; a batch runs it while any
        // system iterates (the others ignore the result).  The V-cycle reads
        // the basis vector and writes Z'_k in FP32.
        float* zk = gm_z.p + ld * k;
        amg->apply_basis<TV>(st, vk, zk, nseg == 1 ? &gm_d.p[0].stop : nullptr);
        spmv_timer.begin(st);
        hipLaunchKernelGGL((k_gm_spmv<TM, float, TV>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st,
                           view_of<TM>(), sg, a, di, (const TV*)vk, (const float*)zk, gm_w.p,
                           gm_partial.p, gm_d.p);
    } else {
        spmv_timer.begin(st);
        hipLaunchKernelGGL((k_gm_spmv<TM, TV, TV>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st,
                           view_of<TM>(), sg, a, di, (const TV*)vk, (const TV*)vk, gm_w.p,
                           gm_partial.p, gm_d.p);
    }
    spmv_timer.end(st);
    hipLaunchKernelGGL(k_gm_dots<TV>, dim3(nbk), dim3(BLOCK), 0, st, sg, gm_w.p, (const TV*)V, ld,
                       k + 1, gm_partial.p, gm_d.p);
    hipLaunchKernelGGL(k_gm_finish_h<TV>, dim3(nseg * GM_MAX), dim3(GMF), 0, st, sg, gm_w.p,
                       (const TV*)V, ld, k + 1, gm_partial.p, gm_d.p);
    hipLaunchKernelGGL(k_gm_orth<TV>, dim3(nbk), dim3(BLOCK), 0, st, sg, (const double*)gm_w.p,
                       V, ld, k + 1, amg ? amg->level0_dinv() : nullptr,
                       amg ? amg->prescaled_input() : nullptr, gm_partial.p, gm_d.p);
    hipLaunchKernelGGL(k_gm_finish_norm, dim3(nseg), dim3(FIN), 0, st, sg, gm_partial.p, gm_d.p,
                       gmf_d.p, m);
    MP_HIP(hipGetLastError());
}

template <typename TM, typename TV>
void PpeSolver::run_gmres(hipStream_t st, const TM* a, const double* di, const double* rhs,
                          double* x, int rel_to_rhs, std::vector<int>& done,
                          std::vector<int>& iters, std::vector<int>& restarts, bool x_zero) {
    const int nbk = nrows / BLOCK;
    const int m = std::max(1, std::min(GM_MAX, restart_m));
    gmres_alloc(st, m);
    const size_t ld = (size_t)nrows;
    TV* V = reinterpret_cast<TV*>(gm_v.p);
    const float* d0 = amg ? amg->level0_dinv() : nullptr;
    float* xs = amg ? amg->prescaled_input() : nullptr;
    MP_HIP(hipMemcpyAsync(ctl_d.p, host_ctl, sizeof(SegCtl) * nseg, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gm_reset, dim3((nseg + 63) / 64), dim3(64), 0, st, nseg, gm_d.p, gmf_d.p,
                       ctl_d.p);
    // V'_0 = rhs - A x, sig[0] = 1 / |rhs - A x| for the systems taking part
    // (mode 0), or for those whose full cycle restarts (mode 1)
    auto start = [&](int mode) {
        hipLaunchKernelGGL((k_gm_residual<TM, TV>), dim3(grid_xcd(nbk)), dim3(BLOCK), 0, st,
                           view_of<TM>(), sg, a, di, rhs, x, V, d0, xs, gm_partial.p, gm_d.p,
                           mode, (mode == 0 && x_zero) ? 1 : 0);
        hipLaunchKernelGGL(k_gm_residual_finish<TV>, dim3(nseg), dim3(FIN), 0, st, sg, rhs, V,
                           gm_partial.p, gm_d.p, gmf_d.p, mode, rel_to_rhs);
        MP_HIP(hipGetLastError());
    };
    start(0);
    const int chunk = std::max(1, check_every);
    GmFlags* hf[2] = {host_gmf, host_gmf + nseg};
    auto any_running = [&](const GmFlags* f) {
        for (int q = 0; q < nseg; ++q)
            if (f[q].done == 0 && !f[q].full) return true;
        return false;
    };
    bool first_cycle = true;
    while (true) {
        // one cycle: chunks of iterations, the flags polled one chunk behind
        // (the first cycle's chunks follow the last run's iteration count)
        int slot = 0;
        bool have_prev = false;
        for (int j = 0; j < m;) {
            int c = chunk;
            if (first_cycle && gm_hint > 0) c = j == 0 ? std::max(1, gm_hint - 1) : 2;
            c = std::min(c, m - j);
            run_gmres_chunk<TM, TV>(st, j, c, m, a, di);
            j += c;
            MP_HIP(hipMemcpyAsync(hf[slot], gmf_d.p, sizeof(GmFlags) * nseg,
                                  hipMemcpyDeviceToHost, st));
            MP_HIP(hipEventRecord(ev[slot], st));
            if (have_prev) {
                MP_HIP(hipEventSynchronize(ev[slot ^ 1]));
                if (!any_running(hf[slot ^ 1])) break;
            }
            have_prev = true;
            slot ^= 1;
        }
        hipLaunchKernelGGL(k_gm_solve, dim3(nseg), dim3(64), 0, st, gm_d.p,
                           gmf_d.p);
        if (amg)
            hipLaunchKernelGGL((k_gm_update_x<float, TV>), dim3(nbk), dim3(BLOCK), 0, st, sg, x,
                               (const float*)gm_z.p, (const TV*)V, ld, gm_d.p);
        else
            hipLaunchKernelGGL((k_gm_update_x<TV, TV>), dim3(nbk), dim3(BLOCK), 0, st, sg, x,
                               (const TV*)V, (const TV*)V, ld, gm_d.p);
        MP_HIP(hipGetLastError());
        MP_HIP(hipMemcpyAsync(host_gmf, gmf_d.p, sizeof(GmFlags) * nseg, hipMemcpyDeviceToHost,
                              st));
        MP_HIP(hipStreamSynchronize(st));
        bool restart = false;
        for (int q = 0; q < nseg; ++q) restart |= host_gmf[q].full && host_gmf[q].done == 0;
        if (!restart) break;
        first_cycle = false;
        start(1);
    }
    int most = 0;
    for (int q = 0; q < nseg; ++q)
        if (host_ctl[q].active) most = std::max(most, host_gmf[q].iter);
    gm_hint = most;
    for (int q = 0; q < nseg; ++q) {
        if (!host_ctl[q].active) continue;
        done[q] = host_gmf[q].done;
        iters[q] = host_gmf[q].iter;
        restarts[q] = host_gmf[q].cycles;
    }
}

void PpeSolver::zero_system(hipStream_t st, double* x, int q) {
    launch_fill(st, x + segs[q].row0, 0.0, segs[q].nv + segs[q].border);
}

SolveStats PpeSolver::solve(hipStream_t st, const double* b_dev, double* x_dev) {
    const int nbk = nrows / BLOCK;
    std::vector<SolveStats> out(nseg);
    // x~ = D x (guess)
    hipLaunchKernelGGL(k_scale, dim3(nbk), dim3(BLOCK), 0, st, nrows, dinv.p, x_dev, xt.p, 1);
    MP_HIP(hipGetLastError());
    std::vector<double> rr0, bb;
    residual64(st, b_dev, rr0, bb);
    for (int q = 0; q < nseg; ++q)
        out[q].guess_rel_res = !(std::isfinite(rr0[q]) && std::isfinite(bb[q])) ? std::nan("")
                               : (bb[q] > 0) ? std::sqrt(rr0[q] / bb[q]) : 0.0;
    if (fallback_x && nseg == 1 && std::isfinite(fallback_rr) &&
        !(rr0[0] <= FALLBACK_FACTOR * FALLBACK_FACTOR * fallback_rr)) {
        // the guess's residual is not finite or far above the second
        // candidate's (the learned guess's linear extrapolation): start from it
        hipLaunchKernelGGL(k_scale, dim3(nbk), dim3(BLOCK), 0, st, nrows, dinv.p, fallback_x,
                           xt.p, 1);
        MP_HIP(hipGetLastError());
        residual64(st, b_dev, rr0, bb);
        out[0].guess_rejected = 1;
    }
    fallback_x = nullptr;
    {
        // a system with a finite right-hand side whose guess is not finite (or
        // leaves a non-finite residual) starts from zero instead
        bool reset = false;
        for (int q = 0; q < nseg; ++q)
            if (std::isfinite(bb[q]) && !std::isfinite(rr0[q])) {
                zero_system(st, xt.p, q);
                out[q].guess_rejected = 1;
                reset = true;
            }
        if (reset) residual64(st, b_dev, rr0, bb);
    }
    std::vector<int> total(nseg, 0), cycles(nseg, 0), rs_total(nseg, 0);
    // the last FP64 residual, when it is still that of the current x~ (the
    // mixed form's defect correction ends with one): the logged residual
    std::vector<double> rr_last;
    std::vector<int> done(nseg, 0), its(nseg, 0), rst(nseg, 0);
    // FP64 BiCGSTAB from the current x~ for the systems in `todo` (budget:
    // iterations each).  BiCGSTAB stops on its recursively updated residual,
    // which drifts from b - A x by rounding; a system whose true residual misses
    // the tolerance restarts from its current x (at most 6 runs).
    // bicg: BiCGSTAB from the start (the mixed form's fallback); otherwise the
    // configured method, a GMRES system switching to BiCGSTAB (from its
    // current x) once it stalls.  Each round runs the GMRES systems, then the
    // BiCGSTAB ones, so every system follows the same sequence as when alone.
    auto fp64_solve = [&](std::vector<char> todo, const std::vector<int>& budget, bool bicg) {
        std::vector<int> used(nseg, 0), runs(nseg, 0);
        std::vector<char> on_bicg(nseg, (bicg || krylov != 1) ? 1 : 0);
        while (true) {
            bool any = false, check = false;
            for (int pass = 0; pass < 2; ++pass) {
                std::vector<char> ran(nseg, 0);
                bool anyp = false;
                for (int q = 0; q < nseg; ++q) {
                    ran[q] = todo[q] && (pass == 0 ? !on_bicg[q] : on_bicg[q]);
                    host_ctl[q].active = ran[q] ? 1 : 0;
                    host_ctl[q].maxit = budget[q] - used[q];
                    host_ctl[q].tol_rel = tol;
                    anyp |= ran[q] != 0;
                }
                if (!anyp) continue;
                any = true;
                if (pass == 0)
                    run_gmres<double, double>(st, aval.p, dinv.p, b_dev, xt.p, 1, done, its, rst);
                else
                    run_bicgstab<double, double>(st, aval.p, dinv.p, b_dev, xt.p, r.p, r0.p, p.p,
                                                 v.p, s.p, t.p, 1, done, its, rst);
                for (int q = 0; q < nseg; ++q) {
                    if (!ran[q]) continue;
                    used[q] += its[q];
                    total[q] += its[q];
                    rs_total[q] += rst[q];
                    cycles[q]++;
                    runs[q]++;
                    if (done[q] == 4) {
                        zero_system(st, xt.p, q);
                        todo[q] = 0;
                    } else if (pass == 0 && (done[q] == 5 || done[q] == 7) &&
                               used[q] < budget[q]) {
                        on_bicg[q] = 1;  // stalled (or broke down): BiCGSTAB next
                        out[q].gmres_switches++;
                    } else if (!(done[q] == 1 || done[q] == 5) || runs[q] >= 6 ||
                               used[q] >= budget[q]) {
                        todo[q] = 0;
                    } else {
                        check = true;
                    }
                }
            }
            if (!any || !check) break;
            std::vector<double> rt, bt;
            residual64(st, b_dev, rt, bt);
            for (int q = 0; q < nseg; ++q)
                if (todo[q] && !(rt[q] > tol * tol * bt[q])) todo[q] = 0;  // converged, or non-finite
        }
    };
    if (precision == 0) {
        fp64_solve(std::vector<char>(nseg, 1), std::vector<int>(nseg, maxit), false);
    } else {
        // FP64 defect correction around an inner BiCGSTAB or GMRES that streams
        // the FP32 copy of the operator while keeping FP64 reductions: each
        // cycle solves A32 e = r with r = b - A x~ from the FP64 operator, to a
        // relative tolerance max(inner_tol, target_margin tol |b| / |r|), then
        // x~ += e.  A system whose cycle breaks down or stalls hands the rest
        // of its solve to FP64 BiCGSTAB.  The projected guess starts near the
        // tolerance (about 1.5 tol), and a solve that ends further below it
        // lets the next step's guess meet it more often: over the 495 timed
        // steps of the 1M-node cavity, margins 0.5 / 0.35 / 0.25 / 0.15 / 0.1
        // gave 7.2 / 6.9 / 6.6 / 6.9 / 7.2 GMRES iterations per step, with
        // 119 / 179 / 224 / 250 / 264 steps needing none.  The other guesses
        // keep 0.5 (with 0.25 the order-4 extrapolation took 16.6 instead of
        // 16.0 iterations per step, and the 64-cavity batch lost 9 %).
        std::vector<double> rr = rr0, rr_prev = rr0, tol2(nseg);
        std::vector<char> act(nseg, 0), fallback(nseg, 0);
        bool undone = false;  // (an undone cycle leaves the residual vector of the undone x~)
        for (int q = 0; q < nseg; ++q) {
            tol2[q] = tol * tol * bb[q];
            act[q] = (rr[q] > tol2[q] && bb[q] > 0) ? 1 : 0;
        }
        while (true) {
            bool any = false;
            for (int q = 0; q < nseg; ++q) {
                act[q] = act[q] && total[q] < maxit;
                host_ctl[q].active = act[q];
                host_ctl[q].maxit = maxit - total[q];
                host_ctl[q].tol_rel = act[q] ? std::max(inner_tol, target_margin * std::sqrt(tol2[q] / rr[q]))
                                             : 1.0;
                any |= act[q] != 0;
            }
            if (!any) break;
            launch_copy(st, rhs64.p, tmp.p, nrows);
            launch_fill(st, e64.p, 0.0, nrows);
            if (krylov == 1)
                run_gmres<float, float>(st, a32k, dinv.p, rhs64.p, e64.p, 1, done, its, rst,
                                        true);
            else
                run_bicgstab<float, double>(st, a32k, dinv.p, rhs64.p, e64.p, r.p, r0.p, p.p,
                                            v.p, s.p, t.p, 1, done, its, rst);
            for (int q = 0; q < nseg; ++q) {
                if (!act[q]) continue;
                total[q] += its[q];
                rs_total[q] += rst[q];
                cycles[q]++;
                if (done[q] == 5 || its[q] == 0) {
                    // breakdown or non-finite values: the correction is dropped
                    fallback[q] = 1;
                    act[q] = 0;
                    zero_system(st, e64.p, q);
                    if (krylov == 1) out[q].gmres_switches++;
                } else if (done[q] == 7) {
                    // a stalled restart cycle: e holds every productive cycle's
                    // correction; the FP64 residual test below keeps or undoes it,
                    // and the rest of the solve runs on FP64 BiCGSTAB
                    fallback[q] = 1;
                    out[q].gmres_switches++;
                }
            }
            // x~ before the correction, kept in rhs64 (free once the inner
            // solve is done); e is zero on the systems that took no part
            launch_copy(st, rhs64.p, xt.p, nrows);
            hipLaunchKernelGGL(k_axpy, dim3(nbk), dim3(BLOCK), 0, st, nrows, 1.0, e64.p, xt.p);
            MP_HIP(hipGetLastError());
            residual64(st, b_dev, rr, bb);
            // a cycle that reduced |r| by less than 10% (or left it non-finite)
            // is undone, and FP64 BiCGSTAB continues from the x~ before it
            std::vector<char> undo(nseg, 0);
            bool any_undo = false;
            for (int q = 0; q < nseg; ++q) {
                if (!act[q]) continue;
                if (!(rr[q] < 0.81 * rr_prev[q])) {
                    fallback[q] = 1;
                    act[q] = 0;
                    undo[q] = 1;
                    any_undo = true;
                    continue;
                }
                rr_prev[q] = rr[q];
                act[q] = rr[q] > tol2[q] && !fallback[q];
                if (!act[q]) fallback[q] = fallback[q] && rr[q] > tol2[q];
            }
            undone |= any_undo;
            if (any_undo)
                for (int q = 0; q < nseg; ++q)
                    if (undo[q]) {
                        const size_t n = (size_t)(segs[q].nv + segs[q].border);
                        MP_HIP(hipMemcpyAsync(xt.p + segs[q].row0, rhs64.p + segs[q].row0,
                                              sizeof(double) * n, hipMemcpyDeviceToDevice, st));
                        rr[q] = rr_prev[q];
                    }
        }
        std::vector<char> fb(nseg, 0);
        std::vector<int> budget(nseg, 0);
        bool anyfb = false;
        for (int q = 0; q < nseg; ++q) {
            fb[q] = fallback[q] && total[q] < maxit;
            budget[q] = maxit - total[q];
            if (fb[q]) {
                fp64_fallbacks++;
                anyfb = true;
            }
        }
        if (anyfb) fp64_solve(fb, budget, true);
        bool zeroed = false;
        for (int q = 0; q < nseg; ++q)
            if (bb[q] == 0) {
                zero_system(st, xt.p, q);
                zeroed = true;
            }
        // rr holds the residual of x~ as it stands (an undone cycle restored
        // both), unless the FP64 fallback or a zero right-hand side moved x~
        if (!anyfb && !zeroed && !undone) rr_last = rr;
    }
    // x = D^-1 x~ and the logged residual |A x - b| (solver.cpp:342-345)
    std::vector<double> rr;
    if (!rr_last.empty())
        rr = rr_last;
    else
        residual64(st, b_dev, rr, bb);
    // the border unknowns came with the last residual (the record of the
    // current x~), so x = D^-1 x~ needs no wait
    std::vector<double> alpha(nseg, 0.0);
    for (int q = 0; q < nseg; ++q)
        if (segs[q].border) alpha[q] = host_sc[q].red[2];
    hipLaunchKernelGGL(k_scale, dim3(nbk), dim3(BLOCK), 0, st, nrows, dinv.p, xt.p, x_dev, 0);
    MP_HIP(hipGetLastError());
    if (spmv_timer.enabled || smooth_timer.enabled) {
        MP_HIP(hipStreamSynchronize(st));
        if (spmv_timer.enabled) spmv_timer.harvest();
        if (smooth_timer.enabled) smooth_timer.harvest();
    }
    SolveStats agg;
    double rr_all = 0, bb_all = 0;
    agg.converged = 1;
    for (int q = 0; q < nseg; ++q) {
        SolveStats& o = out[q];
        o.iterations = total[q];
        o.restarts = rs_total[q];
        o.outer_cycles = cycles[q];
        // a non-finite right-hand side or solution reports NaN, not converged
        const bool finite = std::isfinite(rr[q]) && std::isfinite(bb[q]);
        o.abs_res = std::sqrt(rr[q]);
        o.rel_res = !finite ? std::nan("") : (bb[q] > 0) ? std::sqrt(rr[q] / bb[q]) : 0.0;
        o.converged = (finite && o.rel_res <= tol) ? 1 : 0;
        o.alpha = alpha[q];
        agg.iterations = std::max(agg.iterations, o.iterations);
        agg.restarts += o.restarts;
        agg.outer_cycles = std::max(agg.outer_cycles, o.outer_cycles);
        agg.converged &= o.converged;
        // the worst system; NaN if any is NaN
        agg.rel_res = std::isnan(o.rel_res) ? o.rel_res : std::max(agg.rel_res, o.rel_res);
        agg.guess_rel_res = std::isnan(o.guess_rel_res)
                                ? o.guess_rel_res
                                : std::max(agg.guess_rel_res, o.guess_rel_res);
        rr_all += rr[q];
        bb_all += bb[q];
        total_iterations += o.iterations;
    }
    if (nseg == 1) {
        agg = out[0];
    } else {
        agg.abs_res = std::sqrt(rr_all);
        agg.alpha = 0.0;
    }
    total_solves += 1;
    seg_stats = out;
    return agg;
}

}  // namespace mp
