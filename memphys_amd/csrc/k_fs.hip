// Parts (b) and (d): predictor, pressure right-hand side, corrector, boundary
// fix-up and continuity residual over SELL-64 stencil operators.
//
// Layout: device rows are grouped into slices of 64 (one wavefront); the
// k-th neighbour of slice row l sits at sptr[slice] + 256*(k/4) + 4*l + k%4
// (SELL-64x4), so every group of four neighbour steps is one coalesced
// wavefront load of column ids and of each weight array.  Rows are sorted by stencil length inside windows of
// 4096 rows, so a slice holds rows of (nearly) one length.  One thread owns
// one row and accumulates in stencil order with fma, the oracle's order.
#include "mp_internal.hpp"

namespace mp {

namespace {

__device__ __forceinline__ int row_base(const SellView& sv, int r) {
    return sell_base(sv.sptr, r);
}

// A row's neighbours come in groups of four adjacent slots (SELL-64x4): slot
// s4 = base + 256 g holds neighbours 4g..4g+3, read as one 16-byte load of
// columns and one 32-byte load per weight array.  Entries at or past the
// row's length are skipped, so every sum keeps the order of the stencil.
typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ i4 cols4(const SellView& sv, int s4) {
    return *reinterpret_cast<const i4*>(sv.col + s4);
}
__device__ __forceinline__ d4 vals4(const double* a, int s4) {
    return *reinterpret_cast<const d4*>(a + s4);
}

// Velocity components of one node, interleaved (u, v[, w, 0]): PACK doubles per
// node, so a neighbour's components arrive in one 16- or 32-byte load (one
// cache line) instead of one load from each component array.
template <int DIM>
constexpr int PACK = DIM == 3 ? 4 : 2;
typedef double d2 __attribute__((ext_vector_type(2)));
template <int DIM>
__device__ __forceinline__ d4 packed_at(const double* __restrict__ a, int c) {
    if constexpr (DIM == 3) {
        const d2 t = *reinterpret_cast<const d2*>(a + 4 * (size_t)c);
        return d4{t[0], t[1], a[4 * (size_t)c + 2], 0.0};
    } else {
        const d2 t = *reinterpret_cast<const d2*>(a + 2 * (size_t)c);
        return d4{t[0], t[1], 0.0, 0.0};
    }
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void pack_kernel(int n, const double* __restrict__ u,
                                                     const double* __restrict__ v,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ out) {
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= n) return;
    if constexpr (DIM == 3)
        *reinterpret_cast<d4*>(out + 4 * (size_t)r) = d4{u[r], v[r], w[r], 0.0};
    else
        *reinterpret_cast<d2*>(out + 2 * (size_t)r) = d2{u[r], v[r]};
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void predictor_kernel(
    SellView sv, const double* __restrict__ gx, const double* __restrict__ gy,
    const double* __restrict__ gz, const double* __restrict__ lap,
    const double* __restrict__ upk, double* __restrict__ uh, double* __restrict__ vh,
    double* __restrict__ wh, double* __restrict__ uhpk, double* __restrict__ S_new,
    const double* __restrict__ S_oo, int npad, double rho_all, double mu_all, double dt_all,
    int euler, const int* __restrict__ bprob, const double* __restrict__ prm,
    const double* __restrict__ force) {
    // XCD-aware numbering: each XCD walks a contiguous eighth of the rows, so
    // the neighbour values it gathers stay in its own L2
    const int lb = xcd_block((sv.nv + BLOCK - 1) / BLOCK);
    if (lb < 0) return;
    const int r = lb * BLOCK + threadIdx.x;
    if (r >= sv.nv) return;
    double rho = rho_all, mu = mu_all, dt = dt_all;
    if (bprob) {  // batch: parameters of this block's problem
        const int q = bprob[lb];
        rho = prm[3 * q];
        mu = prm[3 * q + 1];
        dt = prm[3 * q + 2];
    }
    const int base = row_base(sv, r);
    const int len = sv.len[r];
    double acc[DIM][4];
#pragma unroll
    for (int f = 0; f < DIM; ++f)
#pragma unroll
        for (int o = 0; o < 4; ++o) acc[f][o] = 0.0;
    for (int j0 = 0; j0 < len; j0 += 4) {
        const int s4 = base + (j0 << 6);
        const i4 c4 = cols4(sv, s4);
        const d4 A0 = vals4(gx, s4), A1 = vals4(gy, s4), A3 = vals4(lap, s4);
        const d4 A2 = (DIM == 3) ? vals4(gz, s4) : d4{0.0, 0.0, 0.0, 0.0};
        // the four gathers are issued together (slots past the row's length
        // read the row's own node, and their terms are skipped)
        d4 fg[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) fg[e] = packed_at<DIM>(upk, j0 + e < len ? c4[e] : r);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (j0 + e >= len) continue;
            const d4 f = fg[e];
#pragma unroll
            for (int q = 0; q < DIM; ++q) {
                acc[q][0] = fma(A0[e], f[q], acc[q][0]);
                acc[q][1] = fma(A1[e], f[q], acc[q][1]);
                if (DIM == 3) acc[q][2] = fma(A2[e], f[q], acc[q][2]);
                acc[q][3] = fma(A3[e], f[q], acc[q][3]);
            }
        }
    }
    const d4 own = packed_at<DIM>(upk, r);
    double* out[3] = {uh, vh, wh};
    d4 hat = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < DIM; ++q) {
        double conv = -(own[0] * acc[q][0]) - own[1] * acc[q][1];
        if (DIM == 3) conv = conv - own[2] * acc[q][2];
        double S = rho * conv + mu * acc[q][3];
        // body force (fractional_step_1.cpp:104-118): added after the
        // convective and diffusive terms, so it enters the AB2 history too
        if (force) S = S + force[(size_t)q * npad + r];
        S_new[(size_t)q * npad + r] = S;
        double res;
        if (euler)
            res = own[q] + (dt / rho) * S;
        else
            res = own[q] + (dt / rho) * (1.5 * S - 0.5 * S_oo[(size_t)q * npad + r]);
        out[q][r] = res;
        hat[q] = res;
    }
    if constexpr (DIM == 3)
        *reinterpret_cast<d4*>(uhpk + 4 * (size_t)r) = hat;
    else
        *reinterpret_cast<d2*>(uhpk + 2 * (size_t)r) = d2{hat[0], hat[1]};
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void ppe_rhs_kernel(
    SellView sv, const double* __restrict__ gx, const double* __restrict__ gy,
    const double* __restrict__ gz, const double* __restrict__ uh, const double* __restrict__ vh,
    const double* __restrict__ wh, const double* __restrict__ uhpk, const double* __restrict__ u,
    const double* __restrict__ v, const double* __restrict__ w, const double* __restrict__ p_old,
    const uint8_t* __restrict__ bnd, const uint8_t* __restrict__ pdir,
    const double* __restrict__ nrm, int npad, double rho_all, double dt_all,
    double* __restrict__ b, const int* __restrict__ bprob, const double* __restrict__ prm) {
    // XCD-aware numbering: each XCD walks a contiguous eighth of the rows, so
    // the neighbour values it gathers stay in its own L2
    const int lb = xcd_block((sv.nv + BLOCK - 1) / BLOCK);
    if (lb < 0) return;
    const int r = lb * BLOCK + threadIdx.x;
    if (r >= sv.nv) return;  // rows without a stencil (border unknowns, padding) get b = 0
    double rho = rho_all, dt = dt_all;
    if (bprob) {
        const int q = bprob[lb];
        rho = prm[3 * q];
        dt = prm[3 * q + 2];
    }
    double out;
    if (bnd[r]) {
        if (pdir[r]) {
            out = p_old[r];
        } else {
            double s = (u[r] - uh[r]) * nrm[r] + (v[r] - vh[r]) * nrm[(size_t)npad + r];
            if (DIM == 3) s = s + (w[r] - wh[r]) * nrm[2 * (size_t)npad + r];
            out = -rho * s / dt;
        }
    } else {
        const int base = row_base(sv, r);
        const int len = sv.len[r];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j0 = 0; j0 < len; j0 += 4) {
            const int s4 = base + (j0 << 6);
            const i4 c4 = cols4(sv, s4);
            const d4 X = vals4(gx, s4), Y = vals4(gy, s4);
            const d4 Z = (DIM == 3) ? vals4(gz, s4) : d4{0.0, 0.0, 0.0, 0.0};
            d4 fg[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) fg[e] = packed_at<DIM>(uhpk, j0 + e < len ? c4[e] : r);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (j0 + e >= len) continue;
                const d4 f = fg[e];
                ax = fma(X[e], f[0], ax);
                ay = fma(Y[e], f[1], ay);
                if (DIM == 3) az = fma(Z[e], f[2], az);
            }
        }
        out = ax + ay;
        if (DIM == 3) out = out + az;
        out = out * rho / dt;
    }
    b[r] = out;
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void corrector_kernel(
    SellView sv, const double* __restrict__ gx, const double* __restrict__ gy,
    const double* __restrict__ gz, const double* __restrict__ p, const double* __restrict__ uh,
    const double* __restrict__ vh, const double* __restrict__ wh, double* __restrict__ un,
    double* __restrict__ vn, double* __restrict__ wn, double rho_all, double dt_all,
    const int* __restrict__ bprob, const double* __restrict__ prm) {
    // XCD-aware numbering: each XCD walks a contiguous eighth of the rows, so
    // the neighbour values it gathers stay in its own L2
    const int lb = xcd_block((sv.nv + BLOCK - 1) / BLOCK);
    if (lb < 0) return;
    const int r = lb * BLOCK + threadIdx.x;
    if (r >= sv.nv) return;
    double rho = rho_all, dt = dt_all;
    if (bprob) {
        const int q = bprob[lb];
        rho = prm[3 * q];
        dt = prm[3 * q + 2];
    }
    const int base = row_base(sv, r);
    const int len = sv.len[r];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int j0 = 0; j0 < len; j0 += 4) {
        const int s4 = base + (j0 << 6);
        const i4 c4 = cols4(sv, s4);
        const d4 X = vals4(gx, s4), Y = vals4(gy, s4);
        const d4 Z = (DIM == 3) ? vals4(gz, s4) : d4{0.0, 0.0, 0.0, 0.0};
        double pg[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) pg[e] = p[j0 + e < len ? c4[e] : r];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (j0 + e >= len) continue;
            const double pc = pg[e];
            ax = fma(X[e], pc, ax);
            ay = fma(Y[e], pc, ay);
            if (DIM == 3) az = fma(Z[e], pc, az);
        }
    }
    un[r] = uh[r] - (dt * ax) / rho;
    vn[r] = vh[r] - (dt * ay) / rho;
    if (DIM == 3) wn[r] = wh[r] - (dt * az) / rho;
}

// Boundary rows with Neumann pressure: Dirichlet velocity keeps its old value,
// Neumann velocity solves the n.grad row for the diagonal entry
// (fractional_step_1.cpp:159-208).  Boundary stencils hold only interior
// neighbours besides the node itself, so rows are independent.
template <int DIM>
__global__ __launch_bounds__(BLOCK) void boundary_fix_kernel(
    SellView sv, const int* __restrict__ brows, int nb, const double* __restrict__ gx,
    const double* __restrict__ gy, const double* __restrict__ gz, const double* __restrict__ nrm,
    int npad, const uint8_t* __restrict__ pdir, const uint8_t* __restrict__ udir,
    const uint8_t* __restrict__ vdir, const uint8_t* __restrict__ wdir,
    const double* __restrict__ u_old, const double* __restrict__ v_old,
    const double* __restrict__ w_old, double* __restrict__ un, double* __restrict__ vn,
    double* __restrict__ wn) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= nb) return;
    const int r = brows[t];
    if (pdir[r]) return;
    const uint8_t* dirs[3] = {udir, vdir, wdir};
    const double* olds[3] = {u_old, v_old, w_old};
    double* news[3] = {un, vn, wn};
    const double nx = nrm[r], ny = nrm[(size_t)npad + r];
    const double nz = (DIM == 3) ? nrm[2 * (size_t)npad + r] : 0.0;
    const int base = row_base(sv, r);
    const int len = sv.len[r];
    for (int q = 0; q < DIM; ++q) {
        if (dirs[q][r]) {
            news[q][r] = olds[q][r];
        } else {
            double rhs = 0.0, diag = 1.0;
            for (int j0 = 0; j0 < len; j0 += 4) {
                const int s4 = base + (j0 << 6);
                const i4 c4 = cols4(sv, s4);
                const d4 X = vals4(gx, s4), Y = vals4(gy, s4);
                const d4 Z = (DIM == 3) ? vals4(gz, s4) : d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (j0 + e >= len) break;
                    const int c = c4[e];
                    double cf = X[e] * nx + Y[e] * ny;
                    if (DIM == 3) cf = cf + Z[e] * nz;
                    if (c == r)
                        diag = cf;
                    else
                        rhs = rhs - cf * news[q][c];
                }
            }
            news[q][r] = rhs / diag;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void spmv_kernel(SellView sv, const double* __restrict__ val,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= sv.nv) return;
    const int base = row_base(sv, r);
    const int len = sv.len[r];
    double s = 0.0;
    for (int j0 = 0; j0 < len; j0 += 4) {
        const int s4 = base + (j0 << 6);
        const i4 c4 = cols4(sv, s4);
        const d4 A = vals4(val, s4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (j0 + e >= len) break;
            s = fma(A[e], x[c4[e]], s);
        }
    }
    y[r] = s;
}

// partial[blk*4 + {0,1,2,3}] = {max |r| interior, sum |r| interior, max boundary, sum boundary}
template <int DIM>
__global__ __launch_bounds__(BLOCK) void divergence_kernel(
    SellView sv, const double* __restrict__ gx, const double* __restrict__ gy,
    const double* __restrict__ gz, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const uint8_t* __restrict__ bnd, double* __restrict__ partial,
    int nblocks) {
    const int lb = xcd_block(nblocks);  // (XCD-aware, as the predictor)
    if (lb < 0) return;
    const int r = lb * BLOCK + threadIdx.x;
    double vals[4] = {0.0, 0.0, 0.0, 0.0};
    if (r < sv.nv) {
        const int base = row_base(sv, r);
        const int len = sv.len[r];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j0 = 0; j0 < len; j0 += 4) {
            const int s4 = base + (j0 << 6);
            const i4 c4 = cols4(sv, s4);
            const d4 X = vals4(gx, s4), Y = vals4(gy, s4);
            const d4 Z = (DIM == 3) ? vals4(gz, s4) : d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (j0 + e >= len) break;
                const int c = c4[e];
                ax = fma(X[e], u[c], ax);
                ay = fma(Y[e], v[c], ay);
                if (DIM == 3) az = fma(Z[e], w[c], az);
            }
        }
        double res = ax + ay;
        if (DIM == 3) res = res + az;
        const double e = fabs(res);
        if (bnd[r]) {
            vals[2] = e;
            vals[3] = e;
        } else {
            vals[0] = e;
            vals[1] = e;
        }
    }
    __shared__ double sh[4][BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) {
        vals[0] = fmax(vals[0], __shfl_xor(vals[0], off));
        vals[1] += __shfl_xor(vals[1], off);
        vals[2] = fmax(vals[2], __shfl_xor(vals[2], off));
        vals[3] += __shfl_xor(vals[3], off);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 4; ++q) sh[q][wid] = vals[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[4] = {sh[0][0], sh[1][0], sh[2][0], sh[3][0]};
        for (int t = 1; t < BLOCK / 64; ++t) {
            o[0] = fmax(o[0], sh[0][t]);
            o[1] += sh[1][t];
            o[2] = fmax(o[2], sh[2][t]);
            o[3] += sh[3][t];
        }
        for (int q = 0; q < 4; ++q) partial[lb * 4 + q] = o[q];
    }
}

// Implicit operator values in SELL slots (solver.cpp:165-214) and its diagonal.
__global__ __launch_bounds__(BLOCK) void build_operator_kernel(
    SellView sv, int dim, const double* __restrict__ gx, const double* __restrict__ gy,
    const double* __restrict__ gz, const double* __restrict__ lap, const double* __restrict__ nrm,
    int npad, const uint8_t* __restrict__ bnd, const uint8_t* __restrict__ dir, double unsteady,
    double conv, double diff, double* __restrict__ aval, double* __restrict__ diag) {
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= sv.nv) return;
    const int base = row_base(sv, r);
    const int len = sv.len[r];
    double dg = 0.0;
    const bool b = bnd[r] != 0, dr = dir[r] != 0;
    const double n0 = b ? nrm[r] : 0.0, n1 = b ? nrm[(size_t)npad + r] : 0.0;
    const double n2 = (b && dim == 3) ? nrm[2 * (size_t)npad + r] : 0.0;
    for (int j0 = 0; j0 < len; j0 += 4) {
        const int s4 = base + (j0 << 6);
        const i4 c4 = cols4(sv, s4);
        const d4 X = vals4(gx, s4), Y = vals4(gy, s4), L = vals4(lap, s4);
        const d4 Z = (dim == 3) ? vals4(gz, s4) : d4{0.0, 0.0, 0.0, 0.0};
        d4 out = {0.0, 0.0, 0.0, 0.0};  // slots past the row's length stay 0
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (j0 + e >= len) break;
            const int c = c4[e];
            double val;
            if (b) {
                if (dr) {
                    val = (c == r) ? 1.0 : 0.0;
                } else {
                    val = n0 * X[e] + n1 * Y[e];
                    if (dim == 3) val += n2 * Z[e];
                }
            } else {
                val = diff * L[e];
                val += conv * X[e];
                val += conv * Y[e];
                if (dim == 3) val += conv * Z[e];
                if (c == r) val += unsteady;
            }
            out[e] = val;
            if (c == r) dg = val;
        }
        *reinterpret_cast<d4*>(aval + s4) = out;
    }
    diag[r] = dg;
}

}  // namespace

template <int DIM>
void launch_predictor(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                      const double* gz, const double* lap, const double* u, const double* v,
                      const double* w, double* uh, double* vh, double* wh, double* S_new,
                      const double* S_oo, int npad, double rho, double mu, double dt, int euler,
                      double* upk, double* uhpk, const int* bprob, const double* prm,
                      const double* force) {
    const int g = (sv.nv + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(pack_kernel<DIM>, dim3(g), dim3(BLOCK), 0, s, sv.nv, u, v, w, upk);
    hipLaunchKernelGGL(predictor_kernel<DIM>, dim3(grid_xcd(g)), dim3(BLOCK), 0, s, sv, gx, gy, gz, lap, upk,
                       uh, vh, wh, uhpk, S_new, S_oo, npad, rho, mu, dt, euler, bprob, prm, force);
    MP_HIP(hipGetLastError());
}

template <int DIM>
void launch_ppe_rhs(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                    const double* gz, const double* uh, const double* vh, const double* wh,
                    const double* uhpk, const double* u, const double* v, const double* w,
                    const double* p_old, const uint8_t* bnd, const uint8_t* pdir,
                    const double* normal_soa, int npad, double rho, double dt, double* b,
                    const int* bprob, const double* prm) {
    const int g = (sv.nv + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(ppe_rhs_kernel<DIM>, dim3(grid_xcd(g)), dim3(BLOCK), 0, s, sv, gx, gy, gz, uh, vh, wh,
                       uhpk, u, v, w, p_old, bnd, pdir, normal_soa, npad, rho, dt, b, bprob, prm);
    MP_HIP(hipGetLastError());
}

template <int DIM>
void launch_corrector(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                      const double* gz, const double* p, const double* uh, const double* vh,
                      const double* wh, double* un, double* vn, double* wn, int npad, double rho,
                      double dt, const int* bprob, const double* prm) {
    (void)npad;
    const int g = (sv.nv + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(corrector_kernel<DIM>, dim3(grid_xcd(g)), dim3(BLOCK), 0, s, sv, gx, gy, gz, p, uh,
                       vh, wh, un, vn, wn, rho, dt, bprob, prm);
    MP_HIP(hipGetLastError());
}

template <int DIM>
void launch_boundary_fix(hipStream_t s, const SellView& sv, const int* brows, int nb,
                         const double* gx, const double* gy, const double* gz,
                         const double* normal_soa, int npad, const uint8_t* pdir,
                         const uint8_t* udir, const uint8_t* vdir, const uint8_t* wdir,
                         const double* u_old, const double* v_old, const double* w_old,
                         double* un, double* vn, double* wn) {
    if (nb == 0) return;
    const int g = (nb + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(boundary_fix_kernel<DIM>, dim3(g), dim3(BLOCK), 0, s, sv, brows, nb, gx, gy,
                       gz, normal_soa, npad, pdir, udir, vdir, wdir, u_old, v_old, w_old, un, vn,
                       wn);
    MP_HIP(hipGetLastError());
}

void launch_spmv(hipStream_t s, const SellView& sv, const double* val, const double* x,
                 double* y) {
    const int g = (sv.nv + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(spmv_kernel, dim3(g), dim3(BLOCK), 0, s, sv, val, x, y);
    MP_HIP(hipGetLastError());
}

template <int DIM>
void launch_divergence(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                       const double* gz, const double* u, const double* v, const double* w,
                       const uint8_t* bnd, double* partial, int nblocks) {
    hipLaunchKernelGGL(divergence_kernel<DIM>, dim3(grid_xcd(nblocks)), dim3(BLOCK), 0, s, sv, gx, gy, gz, u,
                       v, w, bnd, partial, nblocks);
    MP_HIP(hipGetLastError());
}

void launch_build_operator(hipStream_t s, const SellView& sv, int dim, const double* gx,
                           const double* gy, const double* gz, const double* lap,
                           const double* normal_soa, int npad, const uint8_t* bnd,
                           const uint8_t* dir, double unsteady, double conv, double diff,
                           double* aval, double* diag) {
    const int g = (sv.nv + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(build_operator_kernel, dim3(g), dim3(BLOCK), 0, s, sv, dim, gx, gy, gz, lap,
                       normal_soa, npad, bnd, dir, unsteady, conv, diff, aval, diag);
    MP_HIP(hipGetLastError());
}

#define MP_INST(D)                                                                              \
    template void launch_predictor<D>(hipStream_t, const SellView&, const double*, const double*, \
                                      const double*, const double*, const double*, const double*, \
                                      const double*, double*, double*, double*, double*,        \
                                      const double*, int, double, double, double, int,          \
                                      double*, double*, const int*, const double*,              \
                                      const double*);                                           \
    template void launch_ppe_rhs<D>(hipStream_t, const SellView&, const double*, const double*,   \
                                    const double*, const double*, const double*, const double*,   \
                                    const double*, const double*, const double*, const double*,   \
                                    const double*, const uint8_t*, const uint8_t*, const double*, \
                                    int, double,                                                  \
                                    double, double*, const int*, const double*);                  \
    template void launch_corrector<D>(hipStream_t, const SellView&, const double*, const double*, \
                                      const double*, const double*, const double*, const double*, \
                                      const double*, double*, double*, double*, int, double,      \
                                      double, const int*, const double*);                         \
    template void launch_boundary_fix<D>(hipStream_t, const SellView&, const int*, int,           \
                                         const double*, const double*, const double*,             \
                                         const double*, int, const uint8_t*, const uint8_t*,      \
                                         const uint8_t*, const uint8_t*, const double*,           \
                                         const double*, const double*, double*, double*,          \
                                         double*);                                                \
    template void launch_divergence<D>(hipStream_t, const SellView&, const double*,               \
                                       const double*, const double*, const double*,               \
                                       const double*, const double*, const uint8_t*, double*,     \
                                       int);
MP_INST(2)
MP_INST(3)
#undef MP_INST

}  // namespace mp
