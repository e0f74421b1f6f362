// Reverse Cuthill-McKee order on the GPU (cuthill_mckee_ordering and
// reverse_cuthill_mckee_ordering, general_functions.cpp:807-879, as called by
// CLOUD::re_order_points_reverse_cuthill_mckee, cloud.cpp:315-331), level by
// level.  Sequentially, the queue takes each node in turn, claims its not yet
// visited stencil neighbours in stencil order and appends them sorted by
// (degree, index).  A node reached from level L is therefore claimed by the
// level-L node of lowest queue position whose stencil holds it, and level L+1
// is the list of claimed nodes sorted by (claiming position, degree, index):
// one atomicMin per stencil entry and one radix sort per level give the same
// numbering.  Nodes the start cannot reach restart the search at the
// lowest-index unvisited node of minimum degree.  order[old] = new, reversed.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <stdexcept>
#include <string>
#include <vector>

#include "mp_internal.hpp"

namespace mp {

namespace {

// key of the lowest (degree, index) unvisited node, reduced into *best
__global__ __launch_bounds__(BLOCK) void k_rcm_start(int nv, const int* __restrict__ rp,
                                                     const int* __restrict__ pos, int idx_bits,
                                                     unsigned long long* best) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    unsigned long long k = ~0ull;
    if (i < nv && pos[i] < 0)
        k = ((unsigned long long)(rp[i + 1] - rp[i]) << idx_bits) | (unsigned long long)i;
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = __shfl_xor(k, o);
        k = v < k ? v : k;
    }
    if ((threadIdx.x & 63) == 0 && k != ~0ull) atomicMin(best, k);
}

__global__ void k_rcm_place(const unsigned long long* __restrict__ best, unsigned long long mask,
                            int at, int* __restrict__ numbering, int* __restrict__ pos) {
    const int v = (int)(*best & mask);
    numbering[at] = v;
    pos[v] = at;
}

// every stencil entry of the level's nodes claims its unvisited column for
// the lowest queue position; the first claim appends the column to cand
__global__ __launch_bounds__(BLOCK) void k_rcm_claim(int a, int b, const int* __restrict__ rp,
                                                     const int* __restrict__ col,
                                                     const int* __restrict__ numbering,
                                                     const int* __restrict__ pos,
                                                     int* __restrict__ claim,
                                                     int* __restrict__ cand,
                                                     int* __restrict__ ncand) {
    const int t = a + blockIdx.x * BLOCK + threadIdx.x;
    if (t >= b) return;
    const int v = numbering[t];
    for (int e = rp[v]; e < rp[v + 1]; ++e) {
        const int c = col[e];
        if (pos[c] >= 0) continue;
        if (atomicMin(claim + c, t) == INT_MAX) cand[atomicAdd(ncand, 1)] = c;
    }
}

__global__ __launch_bounds__(BLOCK) void k_rcm_keys(int n, const int* __restrict__ cand,
                                                    const int* __restrict__ claim,
                                                    const int* __restrict__ rp, int idx_bits,
                                                    int deg_bits,
                                                    unsigned long long* __restrict__ keys) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= n) return;
    const int c = cand[t];
    keys[t] = ((unsigned long long)claim[c] << (idx_bits + deg_bits)) |
              ((unsigned long long)(rp[c + 1] - rp[c]) << idx_bits) | (unsigned long long)c;
}

__global__ __launch_bounds__(BLOCK) void k_rcm_append(int n, int at,
                                                      const unsigned long long* __restrict__ keys,
                                                      unsigned long long mask,
                                                      int* __restrict__ numbering,
                                                      int* __restrict__ pos,
                                                      int* __restrict__ claim) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(keys[t] & mask);
    numbering[at + t] = c;
    pos[c] = at + t;
    claim[c] = INT_MAX;
}

__global__ __launch_bounds__(BLOCK) void k_rcm_reverse(int nv, const int* __restrict__ numbering,
                                                       int* __restrict__ order) {
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p < nv) order[numbering[p]] = nv - 1 - p;
}

__global__ __launch_bounds__(BLOCK) void k_fill_int(int* __restrict__ x, int v, int n) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = v;
}

int bits_for(long v) {  // bits to hold 0..v
    int b = 1;
    while ((1L << b) <= v) ++b;
    return b;
}

}  // namespace

void rcm_order(hipStream_t st, int nv, const int* row_ptr, const int* col, int* order,
               int* levels_out) {
    if (nv <= 0) throw std::invalid_argument("rcm: nv must be positive");
    if (row_ptr[0] != 0) throw std::invalid_argument("rcm: row_ptr[0] must be 0");
    int max_deg = 0;
    for (int i = 0; i < nv; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) throw std::invalid_argument("rcm: row_ptr decreases");
        max_deg = std::max(max_deg, row_ptr[i + 1] - row_ptr[i]);
    }
    const long nnz = row_ptr[nv];
    for (long e = 0; e < nnz; ++e)
        if (col[e] < 0 || col[e] >= nv)
            throw std::invalid_argument("rcm: column " + std::to_string(col[e]) +
                                        " out of range");
    const int idx_bits = bits_for(nv - 1);
    const int deg_bits = bits_for(max_deg);
    if (2 * idx_bits + deg_bits > 64)
        throw std::invalid_argument("rcm: graph too large for 64-bit sort keys");
    const unsigned long long mask = (1ull << idx_bits) - 1ull;

    DevBuf<int> drp, dcol, numbering, pos, claim, cand, ncand, dorder;
    DevBuf<unsigned long long> keys, keys_sorted, best;
    drp.alloc((size_t)nv + 1);
    dcol.alloc(std::max<long>(nnz, 1));
    numbering.alloc(nv);
    pos.alloc(nv);
    claim.alloc(nv);
    cand.alloc(nv);
    ncand.alloc(1);
    dorder.alloc(nv);
    keys.alloc(nv);
    keys_sorted.alloc(nv);
    best.alloc(1);
    MP_HIP(hipMemcpyAsync(drp.p, row_ptr, sizeof(int) * (nv + 1), hipMemcpyHostToDevice, st));
    if (nnz > 0)
        MP_HIP(hipMemcpyAsync(dcol.p, col, sizeof(int) * nnz, hipMemcpyHostToDevice, st));
    const int nb = (nv + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_fill_int, dim3(nb), dim3(BLOCK), 0, st, pos.p, -1, nv);
    hipLaunchKernelGGL(k_fill_int, dim3(nb), dim3(BLOCK), 0, st, claim.p, INT_MAX, nv);
    MP_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    MP_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys.p, keys_sorted.p, nv, 0,
                                             2 * idx_bits + deg_bits, st));
    DevBuf<unsigned char> tmp;
    tmp.alloc(std::max<size_t>(tmp_bytes, 1));
    int numbered = 0, levels = 0;
    int* h_n = nullptr;
    MP_HIP(hipHostMalloc((void**)&h_n, sizeof(int), 0));
    try {
        while (numbered < nv) {
            MP_HIP(hipMemsetAsync(best.p, 0xFF, sizeof(unsigned long long), st));
            hipLaunchKernelGGL(k_rcm_start, dim3(nb), dim3(BLOCK), 0, st, nv, drp.p, pos.p,
                               idx_bits, best.p);
            hipLaunchKernelGGL(k_rcm_place, dim3(1), dim3(1), 0, st, best.p, mask, numbered,
                               numbering.p, pos.p);
            MP_HIP(hipGetLastError());
            int a = numbered, b = numbered + 1;
            numbered = b;
            ++levels;
            while (b > a) {
                MP_HIP(hipMemsetAsync(ncand.p, 0, sizeof(int), st));
                hipLaunchKernelGGL(k_rcm_claim, dim3((b - a + BLOCK - 1) / BLOCK), dim3(BLOCK),
                                   0, st, a, b, drp.p, dcol.p, numbering.p, pos.p, claim.p,
                                   cand.p, ncand.p);
                MP_HIP(hipGetLastError());
                MP_HIP(hipMemcpyAsync(h_n, ncand.p, sizeof(int), hipMemcpyDeviceToHost, st));
                MP_HIP(hipStreamSynchronize(st));
                const int n = *h_n;
                if (n == 0) break;
                if (n > nv - numbered) throw std::runtime_error("rcm: level overflow");
                const int gb = (n + BLOCK - 1) / BLOCK;
                hipLaunchKernelGGL(k_rcm_keys, dim3(gb), dim3(BLOCK), 0, st, n, cand.p, claim.p,
                                   drp.p, idx_bits, deg_bits, keys.p);
                MP_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, keys.p,
                                                         keys_sorted.p, n, 0,
                                                         2 * idx_bits + deg_bits, st));
                hipLaunchKernelGGL(k_rcm_append, dim3(gb), dim3(BLOCK), 0, st, n, numbered,
                                   keys_sorted.p, mask, numbering.p, pos.p, claim.p);
                MP_HIP(hipGetLastError());
                a = numbered;
                numbered += n;
                b = numbered;
                ++levels;
            }
        }
        hipLaunchKernelGGL(k_rcm_reverse, dim3(nb), dim3(BLOCK), 0, st, nv, numbering.p,
                           dorder.p);
        MP_HIP(hipGetLastError());
        MP_HIP(hipMemcpyAsync(order, dorder.p, sizeof(int) * nv, hipMemcpyDeviceToHost, st));
        MP_HIP(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipHostFree(h_n);
        throw;
    }
    (void)hipHostFree(h_n);
    if (levels_out) *levels_out = levels;
}

}  // namespace mp
