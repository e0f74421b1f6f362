// Largest-magnitude eigenvalue of a stencil operator on the GPU (replaces
// calc_largest_magnitude_eigenvalue, general_functions.cpp:4-24, which runs
// Spectra's GenEigsSolver<LARGEST_MAGN> with nev = 1, ncv = 10; used by
// PARAMETERS::calc_dt, parameters.cpp:269-296).
//
// Explicitly restarted Arnoldi: m steps of Arnoldi with classical
// Gram-Schmidt applied twice (the products and projections on the device,
// the small sums on the host in a fixed order), eigenvalues of the m x m
// Hessenberg matrix by a shifted complex QR iteration on the host, restart
// from the Ritz vector of the largest-magnitude Ritz value (real and imaginary
// parts added, so a complex-conjugate pair keeps its two-dimensional invariant
// subspace), until the Ritz residual |h(m+1,m) y_m| / |lambda| <= tol.
#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <vector>

#include "eig_host.hpp"
#include "mp_internal.hpp"

namespace mp {

namespace {

using mp_eig::cplx;
using mp_eig::hessenberg_eigenvalues;
using mp_eig::hessenberg_eigenvector;

// partial[j * nb + b] = sum over block b's rows of V_j . w  (j < k)
__global__ __launch_bounds__(BLOCK) void k_multidot(int n, int k, const double* __restrict__ V,
                                                    size_t ld, const double* __restrict__ w,
                                                    double* __restrict__ partial) {
    __shared__ double sh[BLOCK / 64];
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const double wi = i < n ? w[i] : 0.0;
    for (int j = 0; j < k; ++j) {
        double v = i < n ? V[(size_t)j * ld + i] * wi : 0.0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = sh[0];
            for (int t = 1; t < BLOCK / 64; ++t) s += sh[t];
            partial[(size_t)j * gridDim.x + blockIdx.x] = s;
        }
        __syncthreads();
    }
}

// w = a w - sum_j c_j V_j
__global__ __launch_bounds__(BLOCK) void k_combine(int n, int k, const double* __restrict__ V,
                                                   size_t ld, const double* __restrict__ c,
                                                   double a, double* __restrict__ w) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double s = a * w[i];
    for (int j = 0; j < k; ++j) s -= c[j] * V[(size_t)j * ld + i];
    w[i] = s;
}

__global__ __launch_bounds__(BLOCK) void k_start(int n, double* __restrict__ v) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    // deterministic start with no special structure: a hashed pattern in [0.5, 1.5)
    unsigned int h = (unsigned int)i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    v[i] = 0.5 + (double)(h & 0xFFFFFF) / 16777216.0;
}

}  // namespace

// Largest-magnitude eigenvalue of the operator with slot values `val` over
// the rows of sv (padding rows contribute zero rows).  Returns the number of
// Arnoldi restarts used; throws when it does not converge.
int largest_eigenvalue(hipStream_t st, const SellView& sv, const double* val, int m, double tol,
                       int max_restarts, double* re, double* im) {
    const int n = sv.nv;
    const size_t ld = (size_t)n + 64;
    const int nb = (n + BLOCK - 1) / BLOCK;
    DevBuf<double> V, w, part, coef;
    V.alloc(ld * (m + 1));
    w.alloc(ld);
    part.alloc((size_t)nb * (m + 1));
    coef.alloc(m + 1);
    MP_HIP(hipMemsetAsync(V.p, 0, sizeof(double) * V.n, st));
    hipLaunchKernelGGL(k_start, dim3(nb), dim3(BLOCK), 0, st, n, V.p);
    std::vector<double> hp((size_t)nb * (m + 1)), c(m + 1);
    // projections of w on V_0..V_{k-1}, summed on the host in block order
    auto dots = [&](int k, const double* vec) {
        hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(BLOCK), 0, st, n, k, V.p, ld, vec, part.p);
        MP_HIP(hipGetLastError());
        MP_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * (size_t)nb * k,
                              hipMemcpyDeviceToHost, st));
        MP_HIP(hipStreamSynchronize(st));
        for (int j = 0; j < k; ++j) {
            double s = 0.0;
            for (int b = 0; b < nb; ++b) s += hp[(size_t)j * nb + b];
            c[j] = s;
        }
    };
    auto combine = [&](int k, double a, double* vec) {
        MP_HIP(hipMemcpyAsync(coef.p, c.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_combine, dim3(nb), dim3(BLOCK), 0, st, n, k, V.p, ld, coef.p, a, vec);
        MP_HIP(hipGetLastError());
    };
    // normalise the start vector (V_0 <- V_0 / |V_0|, through w)
    auto normalise_into = [&](double* src, double* dst) -> double {
        MP_HIP(hipMemcpyAsync(V.p + (size_t)m * ld, src, sizeof(double) * n,
                              hipMemcpyDeviceToDevice, st));
        // |src|^2 as the projection of src on itself (stored in slot m)
        hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(BLOCK), 0, st, n, 1, V.p + (size_t)m * ld,
                           ld, src, part.p);
        MP_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
        MP_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += hp[b];
        const double nrm = std::sqrt(s);
        c[0] = 0.0;
        MP_HIP(hipMemcpyAsync(coef.p, c.data(), sizeof(double), hipMemcpyHostToDevice, st));
        if (dst != src)
            MP_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_combine, dim3(nb), dim3(BLOCK), 0, st, n, 0, V.p, ld, coef.p,
                           nrm > 0 ? 1.0 / nrm : 0.0, dst);
        MP_HIP(hipGetLastError());
        return nrm;
    };
    normalise_into(V.p, V.p);
    std::vector<double> H((size_t)(m + 1) * m);
    for (int restart = 0; restart <= max_restarts; ++restart) {
        std::fill(H.begin(), H.end(), 0.0);
        int k_used = m;
        for (int j = 0; j < m; ++j) {
            launch_spmv(st, sv, val, V.p + (size_t)j * ld, w.p);
            // classical Gram-Schmidt, twice
            for (int pass = 0; pass < 2; ++pass) {
                dots(j + 1, w.p);
                for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] += c[i];
                combine(j + 1, 1.0, w.p);
            }
            const double hn = normalise_into(w.p, V.p + (size_t)(j + 1) * ld);
            H[(size_t)(j + 1) * m + j] = hn;
            if (hn == 0.0) {  // invariant subspace: the Ritz values are exact
                k_used = j + 1;
                break;
            }
        }
        std::vector<double> Hm((size_t)k_used * k_used);
        for (int r = 0; r < k_used; ++r)
            for (int q = 0; q < k_used; ++q) Hm[(size_t)r * k_used + q] = H[(size_t)r * m + q];
        std::vector<cplx> ev = hessenberg_eigenvalues(Hm, k_used);
        int best = 0;
        for (int q = 1; q < k_used; ++q)
            if (std::abs(ev[q]) > std::abs(ev[best]) ||
                (std::abs(ev[q]) == std::abs(ev[best]) && ev[q].imag() > ev[best].imag()))
                best = q;
        const cplx lam = ev[best];
        std::vector<cplx> y = hessenberg_eigenvector(Hm, k_used, lam);
        const double hnext = (k_used == m) ? H[(size_t)m * m + (m - 1)] : 0.0;
        const double resid = std::abs(hnext) * std::abs(y[k_used - 1]) /
                             std::max(std::abs(lam), 1e-300);
        *re = lam.real();
        *im = std::abs(lam.imag());
        if (resid <= tol || k_used < m) return restart;
        // restart vector: V (Re y + Im y)
        for (int q = 0; q < k_used; ++q) c[q] = -(y[q].real() + y[q].imag());
        launch_fill(st, w.p, 0.0, n);
        combine(k_used, 1.0, w.p);
        normalise_into(w.p, V.p);
    }
    throw std::runtime_error("largest_eigenvalue: Arnoldi did not converge");
}

}  // namespace mp
