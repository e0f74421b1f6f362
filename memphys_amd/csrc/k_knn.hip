// Stencil selection on the GPU (CLOUD::calc_cloud_points_fast,
// cloud.cpp:570-648):
//   interior node: its k nearest nodes among all nodes, itself first;
//   boundary node: itself, then its k-1 nearest interior nodes.
// Neighbours are ordered by (squared distance, index): the k smallest pairs of
// that total order, which is what the reference's kd-tree returns up to its
// traversal order on exact ties (broken here by the lower index).
//
// A uniform grid of cells (about two nodes per cell) is built on the device:
// cell ids, counts, an exclusive scan and a scatter.  One wavefront serves one
// node: it visits the cells ring by ring (Chebyshev distance r around the
// node's cell, 64 cells a batch), compacts the candidates of each batch into
// chunks of up to 64, and merges every chunk into a sorted top-k list in LDS
// by rank counting.  After ring r every unvisited node lies at least r cell
// widths away, so the search stops once the k-th distance is below that.
// The result does not depend on the order of nodes inside a cell.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "mp_internal.hpp"

namespace mp {

namespace {

struct KnnGrid {
    double lo[3];
    double cell;
    int nc[3];
    int dim;
};

__host__ __device__ inline int cell_axis(double x, double lo, double cell, int nc) {
    const int c = (int)floor((x - lo) / cell);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ __launch_bounds__(BLOCK) void k_cell_ids(int nv, const double* __restrict__ xyz,
                                                    KnnGrid g, int* __restrict__ cid,
                                                    int* __restrict__ cnt) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nv) return;
    int lin = 0, mul = 1;
    for (int a = 0; a < g.dim; ++a) {
        lin += cell_axis(xyz[(size_t)i * g.dim + a], g.lo[a], g.cell, g.nc[a]) * mul;
        mul *= g.nc[a];
    }
    cid[i] = lin;
    atomicAdd(cnt + lin, 1);
}

__global__ __launch_bounds__(BLOCK) void k_cell_fill(int nv, const int* __restrict__ cid,
                                                     int* __restrict__ fill,
                                                     int* __restrict__ items) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nv) return;
    items[atomicAdd(fill + cid[i], 1)] = i;
}

__device__ __forceinline__ bool key_less(double da, int ia, double db, int ib) {
    return da < db || (da == db && ia < ib);
}

// One node per block of one wavefront (so that __syncthreads is a wave barrier).
// KCAP: capacity of the sorted candidate lists (the largest stencil).
template <int KCAP>
__global__ __launch_bounds__(64) void k_knn(int nv, const double* __restrict__ xyz,
                                            const uint8_t* __restrict__ bnd,
                                            const int* __restrict__ kpn,
                                            const int* __restrict__ row_ptr, KnnGrid g,
                                            const int* __restrict__ cstart,
                                            const int* __restrict__ items,
                                            int* __restrict__ col_out) {
    __shared__ double td[2][KCAP];
    __shared__ int ti[2][KCAP];
    __shared__ double cd[64];
    __shared__ int cix[64];
    __shared__ int cell_first[64];
    __shared__ int pref[65];
    const int q = blockIdx.x;
    const int lane = threadIdx.x;
    const int dim = g.dim;
    const bool qb = bnd[q] != 0;
    const int need = qb ? kpn[q] - 1 : kpn[q];
    double qx[3] = {0.0, 0.0, 0.0};
    int cq[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        qx[a] = xyz[(size_t)q * dim + a];
        cq[a] = cell_axis(qx[a], g.lo[a], g.cell, g.nc[a]);
    }
    int max_ring = 0;
    for (int a = 0; a < dim; ++a) max_ring = max(max_ring, g.nc[a]);
    int nT = 0, cur = 0;
    for (int r = 0; need > 0; ++r) {
        const int side = 2 * r + 1;
        const long ncube = (dim == 3) ? (long)side * side * side : (long)side * side;
        for (long base = 0; base < ncube; base += 64) {
            const long t = base + lane;
            int st = 0, cnt = 0;
            if (t < ncube) {
                const int off[3] = {(int)(t % side) - r, (int)((t / side) % side) - r,
                                    (dim == 3) ? (int)(t / ((long)side * side)) - r : 0};
                int cheb = 0, lin = 0, mul = 1;
                bool inside = true;
                for (int a = 0; a < dim; ++a) {
                    const int c = cq[a] + off[a];
                    cheb = max(cheb, abs(off[a]));
                    inside = inside && c >= 0 && c < g.nc[a];
                    lin += c * mul;
                    mul *= g.nc[a];
                }
                if (cheb == r && inside) {
                    st = cstart[lin];
                    cnt = cstart[lin + 1] - st;
                }
            }
            // inclusive prefix of the batch's item counts across the wavefront
            int incl = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            const int total = __shfl(incl, 63);
            cell_first[lane] = st;
            pref[lane + 1] = incl;
            if (lane == 0) pref[0] = 0;
            __syncthreads();
            for (int b0 = 0; b0 < total; b0 += 64) {
                const int j = b0 + lane;
                bool valid = j < total;
                double d2 = 0.0;
                int id = 0;
                if (valid) {
                    int lo = 0, hi = 63;  // cell of item j: pref[lo] <= j < pref[lo + 1]
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (pref[mid] <= j)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    id = items[cell_first[lo] + (j - pref[lo])];
                    if (qb && bnd[id]) {
                        valid = false;
                    } else {
                        for (int a = 0; a < dim; ++a) {
                            const double d = xyz[(size_t)id * dim + a] - qx[a];
                            d2 = d2 + d * d;
                        }
                        if (nT == need &&
                            !key_less(d2, id, td[cur][need - 1], ti[cur][need - 1]))
                            valid = false;
                    }
                }
                const unsigned long long m = __ballot(valid);
                const int nc = __popcll(m);
                if (nc == 0) continue;
                if (valid) {
                    const int pos = __popcll(m & ((1ull << lane) - 1ull));
                    cd[pos] = d2;
                    cix[pos] = id;
                }
                __syncthreads();
                const int nxt = cur ^ 1;
                const int n_new = min(need, nT + nc);
                if (valid) {
                    int lo = 0, hi = nT;  // T entries below the candidate
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (key_less(td[cur][mid], ti[cur][mid], d2, id))
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    int p = lo;
                    for (int u = 0; u < nc; ++u) p += key_less(cd[u], cix[u], d2, id) ? 1 : 0;
                    if (p < n_new) {
                        td[nxt][p] = d2;
                        ti[nxt][p] = id;
                    }
                }
                for (int i = lane; i < nT; i += 64) {
                    const double dv = td[cur][i];
                    const int iv = ti[cur][i];
                    int p = i;
                    for (int u = 0; u < nc; ++u) p += key_less(cd[u], cix[u], dv, iv) ? 1 : 0;
                    if (p < n_new) {
                        td[nxt][p] = dv;
                        ti[nxt][p] = iv;
                    }
                }
                __syncthreads();
                cur = nxt;
                nT = n_new;
            }
            __syncthreads();
        }
        if (nT == need) {
            // every unvisited node lies in a ring >= r + 1, at least r cells away
            // (a hair less is assumed, against rounding in the cell index)
            const double bound = (double)r * g.cell * (1.0 - 1e-9);
            if (td[cur][need - 1] < bound * bound) break;
        }
        if (r >= max_ring) break;
    }
    int* out = col_out + row_ptr[q];
    if (qb) {
        if (lane == 0) out[0] = q;
        for (int i = lane; i < nT; i += 64) out[1 + i] = ti[cur][i];
        return;
    }
    // the node itself leads its stencil; an exact duplicate point of lower
    // index is moved behind it (if k duplicates coincide, the node replaces
    // the last of them)
    int self = need;
    for (int i = lane; i < nT; i += 64)
        if (ti[cur][i] == q) self = i;
    for (int o = 32; o > 0; o >>= 1) self = min(self, __shfl_xor(self, o));
    if (lane == 0) out[0] = q;
    for (int i = lane; i < nT; i += 64) {
        if (i < self)
            out[1 + i] = ti[cur][i];
        else if (i > self)
            out[i] = ti[cur][i];
    }
}

}  // namespace

void knn_stencils(hipStream_t st, int nv, int dim, const double* xyz, const uint8_t* boundary,
                  const int* k_per_node, int* row_ptr, int* col) {
    if (nv <= 0 || (dim != 2 && dim != 3)) throw std::invalid_argument("knn: bad nv or dim");
    long n_int = 0;
    int kmax = 0;
    for (int i = 0; i < nv; ++i) {
        n_int += boundary[i] ? 0 : 1;
        kmax = std::max(kmax, k_per_node[i]);
    }
    row_ptr[0] = 0;
    for (int i = 0; i < nv; ++i) {
        const int k = k_per_node[i];
        const long need = boundary[i] ? k - 1 : k;
        const long avail = boundary[i] ? n_int : nv;
        if (k < 1 || k > MAX_STENCIL || need > avail)
            throw std::invalid_argument("knn: stencil size " + std::to_string(k) + " of node " +
                                        std::to_string(i) + " is outside [1, " +
                                        std::to_string(MAX_STENCIL) +
                                        "] or exceeds the available nodes");
        row_ptr[i + 1] = row_ptr[i] + k;
    }
    KnnGrid g{};
    g.dim = dim;
    double hi[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        g.lo[a] = hi[a] = xyz[a];
        for (int i = 0; i < nv; ++i) {
            g.lo[a] = std::min(g.lo[a], xyz[(size_t)i * dim + a]);
            hi[a] = std::max(hi[a], xyz[(size_t)i * dim + a]);
        }
    }
    double vol = 1.0;
    for (int a = 0; a < dim; ++a) vol *= std::max(hi[a] - g.lo[a], 1e-300);
    g.cell = std::pow(2.0 * vol / (double)nv, 1.0 / dim);  // about two nodes per cell
    if (!(g.cell > 0.0) || !std::isfinite(g.cell)) g.cell = 1.0;
    long ncell = 0;
    for (;;) {
        ncell = 1;
        for (int a = 0; a < dim; ++a) {
            g.nc[a] = std::max(1, (int)std::ceil((hi[a] - g.lo[a]) / g.cell));
            ncell *= g.nc[a];
        }
        if (ncell <= 4L * nv + 64) break;  // degenerate extents: coarser cells
        g.cell *= 1.5;
    }
    for (int a = dim; a < 3; ++a) g.nc[a] = 1;

    DevBuf<double> dxyz;
    DevBuf<uint8_t> dbnd;
    DevBuf<int> dk, drp, dcol, cid, cnt, cstart, fill, items;
    dxyz.alloc((size_t)nv * dim);
    dbnd.alloc(nv);
    dk.alloc(nv);
    drp.alloc((size_t)nv + 1);
    dcol.alloc(row_ptr[nv]);
    cid.alloc(nv);
    cnt.alloc(ncell + 1);
    cstart.alloc(ncell + 1);
    fill.alloc(ncell + 1);
    items.alloc(nv);
    MP_HIP(hipMemcpyAsync(dxyz.p, xyz, sizeof(double) * nv * dim, hipMemcpyHostToDevice, st));
    MP_HIP(hipMemcpyAsync(dbnd.p, boundary, nv, hipMemcpyHostToDevice, st));
    MP_HIP(hipMemcpyAsync(dk.p, k_per_node, sizeof(int) * nv, hipMemcpyHostToDevice, st));
    MP_HIP(hipMemcpyAsync(drp.p, row_ptr, sizeof(int) * (nv + 1), hipMemcpyHostToDevice, st));
    MP_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * (ncell + 1), st));
    const int nb = (nv + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_cell_ids, dim3(nb), dim3(BLOCK), 0, st, nv, dxyz.p, g, cid.p, cnt.p);
    MP_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    MP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt.p, cstart.p, (int)ncell + 1,
                                            st));
    DevBuf<unsigned char> tmp;
    tmp.alloc(std::max<size_t>(tmp_bytes, 1));
    MP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, cnt.p, cstart.p, (int)ncell + 1,
                                            st));
    MP_HIP(hipMemcpyAsync(fill.p, cstart.p, sizeof(int) * (ncell + 1), hipMemcpyDeviceToDevice,
                          st));
    hipLaunchKernelGGL(k_cell_fill, dim3(nb), dim3(BLOCK), 0, st, nv, cid.p, fill.p, items.p);
    MP_HIP(hipGetLastError());
    MP_HIP(hipMemsetAsync(dcol.p, 0xFF, sizeof(int) * row_ptr[nv], st));  // -1: unfilled
    // (the small lists keep the common stencil sizes at full occupancy)
    if (kmax <= 96)
        hipLaunchKernelGGL(k_knn<96>, dim3(nv), dim3(64), 0, st, nv, dxyz.p, dbnd.p, dk.p, drp.p,
                           g, cstart.p, items.p, dcol.p);
    else
        hipLaunchKernelGGL(k_knn<MAX_STENCIL>, dim3(nv), dim3(64), 0, st, nv, dxyz.p, dbnd.p,
                           dk.p, drp.p, g, cstart.p, items.p, dcol.p);
    MP_HIP(hipGetLastError());
    MP_HIP(hipMemcpyAsync(col, dcol.p, sizeof(int) * row_ptr[nv], hipMemcpyDeviceToHost, st));
    MP_HIP(hipStreamSynchronize(st));
}

}  // namespace mp
