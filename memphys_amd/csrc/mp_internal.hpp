// Internal declarations shared by the HIP translation units of libmemphys_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/memphys_hip.h"

namespace mp {

constexpr int SLICE = 64;       // rows per SELL slice (one wavefront)
constexpr int SGROUP = 4;       // neighbours of a row stored side by side in a slice
constexpr int SIGMA = 4096;     // rows per length-sorting window
constexpr int BLOCK = 256;      // threads per block of the row kernels
// Largest stencil (k) and monomial count (m) accepted: the reference's
// parameter range (poly_deg up to 15, parameters.cpp:84-88, 122-130) gives
// m = 136 in 2D; stencils of M = k + m above what LDS holds run from a global
// work area.  The register-resident weights kernels and the mixed kernel see
// m <= WAVE_MAX_POLY only (their stencils have M <= 96, and M >= 3m).
constexpr int MAX_STENCIL = 512;
constexpr int MAX_POLY = 136;
constexpr int WAVE_MAX_POLY = 35;
// dynamic LDS the weights kernel's per-stencil system may use (of 160 KB per CU,
// leaving room for its static arrays)
constexpr size_t WEIGHTS_LDS_MAX = 150 * 1024;

struct HipError {
    hipError_t code;
    const char* what;
};

#define MP_HIP(call)                                         \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) throw ::mp::HipError{_e, #call}; \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        free();
        if (count == 0) count = 1;
        MP_HIP(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            free();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
};

// Number of XCDs of the MI355X: the dispatcher hands workgroup b to XCD b % 8.
constexpr int NXCD = 8;

// XCD-aware block numbering: launch grid_xcd(G) blocks; block b works on
// logical block (b % 8) * ceil(G/8) + b / 8, so each XCD (and its L2) owns a
// contiguous range of rows.  Returns -1 for the padding blocks.
__host__ __device__ inline int grid_xcd(int g) { return ((g + NXCD - 1) / NXCD) * NXCD; }
__device__ __forceinline__ int xcd_block(int g) {
    const int per = (g + NXCD - 1) / NXCD;
    const int lb = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
    return lb < g ? lb : -1;
}

// SELL-64x4 slot numbering: slice s holds rows 64s..64s+63; neighbour j of
// slice row l sits at sptr[s] + 256*(j/4) + 4*l + j%4, so each lane reads its
// row's neighbours four at a time (one 16-byte load of columns or FP32
// values) and the wavefront still covers 256 consecutive slots per group.
// Slice widths are padded to a multiple of 4; padding slots hold the row's
// own column and value 0.
__host__ __device__ inline int sell_base(const int* sptr, int r) {
    return sptr[r >> 6] + ((r & 63) << 2);
}
__host__ __device__ inline int sell_off(int j) { return ((j >> 2) << 8) + (j & 3); }
__host__ __device__ inline int sell_width4(int w) { return (w + 3) & ~3; }

template <typename T>
struct Vec4 {
    typedef T type __attribute__((ext_vector_type(4)));
};

// Compressed columns: a slice whose columns fall into at most CW_COUNT
// windows of 2^CW_BITS consecutive ids stores each slot's column in 16 bits,
// (window << CW_BITS) | (column - base[window]), with the slice's window bases
// in cbase[CW_COUNT * s ...]; cbase[CW_COUNT * s] < 0 marks a slice that keeps
// its 32-bit columns.  The decode is exact, so either form gives the same sums.
constexpr int CW_BITS = 13;
constexpr int CW_COUNT = 8;

__device__ __forceinline__ int col_decode(uint32_t e, const int (&b)[CW_COUNT]) {
    const uint32_t w = e >> CW_BITS;
    const int t0 = (w & 1u) ? b[1] : b[0];
    const int t1 = (w & 1u) ? b[3] : b[2];
    const int t2 = (w & 1u) ? b[5] : b[4];
    const int t3 = (w & 1u) ? b[7] : b[6];
    const int u0 = (w & 2u) ? t1 : t0;
    const int u1 = (w & 2u) ? t3 : t2;
    return ((w & 4u) ? u1 : u0) + (int)(e & ((1u << CW_BITS) - 1u));
}

template <typename V>
__device__ __forceinline__ V load_values(const V* p) {
    if constexpr (sizeof(V) > 16)
        return *p;
    else
        return __builtin_nontemporal_load(p);
}

// The neighbour loop of one slice row (groups of four slots from slot b0 +
// 4 lane, w4 groups): load(g) fetches group g's stored columns and decode()
// turns them into column ids.  The stored columns and the values of U4
// groups are all requested before the first decode, so their loads overlap.
template <int U4, typename T, typename TM, typename LD, typename DC, typename FV, typename FX>
__device__ __forceinline__ T sell4_loop(int b0, int w4, int lane, const TM* __restrict__ a,
                                        LD load, DC decode, FV fv, FX fx) {
    typedef typename Vec4<TM>::type v4;
    typedef decltype(load(0)) raw4;
    const v4* ap = reinterpret_cast<const v4*>(a + b0) + lane;
    T acc = (T)0;
    int g = 0;
    for (; g + U4 <= w4; g += U4) {
        raw4 e[U4];
        v4 v[U4];
        T xv[4 * U4];
#pragma unroll
        for (int u = 0; u < U4; ++u) e[u] = load(g + u);
#pragma unroll
        for (int u = 0; u < U4; ++u) v[u] = load_values(ap + (g + u) * SLICE);
#pragma unroll
        for (int u = 0; u < U4; ++u) {
            const typename Vec4<int>::type c = decode(e[u]);
            xv[4 * u + 0] = fx(c.x);
            xv[4 * u + 1] = fx(c.y);
            xv[4 * u + 2] = fx(c.z);
            xv[4 * u + 3] = fx(c.w);
        }
#pragma unroll
        for (int u = 0; u < U4; ++u) {
            acc = fma(fv(v[u].x), xv[4 * u + 0], acc);
            acc = fma(fv(v[u].y), xv[4 * u + 1], acc);
            acc = fma(fv(v[u].z), xv[4 * u + 2], acc);
            acc = fma(fv(v[u].w), xv[4 * u + 3], acc);
        }
    }
    for (; g < w4; ++g) {
        const raw4 e = load(g);
        const v4 v = load_values(ap + g * SLICE);
        const typename Vec4<int>::type c = decode(e);
        const T x0 = fx(c.x), x1 = fx(c.y), x2 = fx(c.z), x3 = fx(c.w);
        acc = fma(fv(v.x), x0, acc);
        acc = fma(fv(v.y), x1, acc);
        acc = fma(fv(v.z), x2, acc);
        acc = fma(fv(v.w), x3, acc);
    }
    return acc;
}

// Sum over the neighbours of one slice row, in neighbour order, of
// fv(value) * fx(column), as an fma chain: slice s, lane = slice row.  The
// four neighbours of a group come in one vector load of columns (16-bit
// compressed ones when col16 is given and the slice has them) and one of
// values; U4 groups are loaded before their gathers.  Columns and values up
// to 4 bytes use non-temporal loads so the gathered vector stays cached; FP64
// values (32 bytes a lane, fetched as two 16-byte halves) use ordinary loads,
// which keep the line in cache between the halves.
template <int U4, typename T, typename TM, typename FV, typename FX>
__device__ __forceinline__ T sell4_accumulate(const int* __restrict__ sptr,
                                              const int* __restrict__ col,
                                              const uint16_t* __restrict__ col16,
                                              const int* __restrict__ cbase,
                                              const TM* __restrict__ a, int s, int lane, FV fv,
                                              FX fx) {
    typedef typename Vec4<int>::type c4;
    typedef typename Vec4<uint16_t>::type e4;
    const int b0 = sptr[s];
    const int w4 = (sptr[s + 1] - b0) >> 8;
    if (col16 != nullptr) {
        const int* bp = cbase + CW_COUNT * s;
        const int base[CW_COUNT] = {bp[0], bp[1], bp[2], bp[3], bp[4], bp[5], bp[6], bp[7]};
        if (base[0] >= 0) {
            const e4* ep = reinterpret_cast<const e4*>(col16 + b0) + lane;
            return sell4_loop<U4, T>(
                b0, w4, lane, a, [&](int g) { return __builtin_nontemporal_load(ep + g * SLICE); },
                [&](e4 e) {
                    c4 c;
                    c.x = col_decode(e.x, base);
                    c.y = col_decode(e.y, base);
                    c.z = col_decode(e.z, base);
                    c.w = col_decode(e.w, base);
                    return c;
                },
                fv, fx);
        }
    }
    const c4* cp = reinterpret_cast<const c4*>(col + b0) + lane;
    return sell4_loop<U4, T>(
        b0, w4, lane, a, [&](int g) { return __builtin_nontemporal_load(cp + g * SLICE); },
        [](c4 c) { return c; }, fv, fx);
}

// SELL-64x4 view of the stencil set in device order.
struct SellView {
    int nv;            // real rows
    int nslices;
    const int* sptr;   // [nslices+1] first slot of each slice
    const int* col;    // [slots] device column ids
    const int* len;    // [nslices*64] row lengths (0 on ghost rows)
    const uint16_t* col16 = nullptr;  // [slots] compressed column ids (or null)
    const int* cbase = nullptr;       // [nslices*CW_COUNT] window bases
};

// Device-side Krylov scalars (FP64) of one system.  Written only by the
// single-block finishing kernels, read by every block of the following kernels.
struct KScalars {
    double rho, rho_old, alpha, omega, beta;
    double r0sq, bb, rr, tol2, eps2, tol_rel;
    double red[4];  // results of the last residual reduction
    // 0 running, 1 converged, 2 iteration cap, 3 restart needed, 4 zero
    // right-hand side, 5 breakdown / non-finite, 6 not part of this solve
    int done;
    int iter, maxit, restarts;
};

// One independent system of a batch: real rows [row0, row0 + nv), its border
// unknown (full-Neumann Lagrange multiplier) at row row0 + nv when border != 0,
// and the row blocks [blk0, blk0 + nblk) of BLOCK rows it owns.  Systems start
// on block boundaries, so a block never mixes two systems.
struct Seg {
    int row0, nv, blk0, nblk, border;
};

struct SegView {
    int nseg, nrows;    // systems; rows of all systems incl. padding (multiple of BLOCK)
    const Seg* seg;     // [nseg]
    const int* bseg;    // [nrows / BLOCK] system of each row block
};

// Per-system controls of one Krylov run (uploaded before it starts).
struct SegCtl {
    int active, maxit;
    double tol_rel;
};

// ---- launchers (defined in the .hip files) ----
// rows == nullptr: every row; otherwise the listed device rows.
// FP64 weights of the device rows `rows` (stencil lengths `lens`): the
// register-resident kernel for systems of M = k + m <= 96 (one or two
// wavefronts per stencil), the LDS kernel beyond.  dev_rows receives the
// grouped row list and must stay allocated until the launches complete.
constexpr int WAVE_CLASSES = 5;
void launch_weights_fp64_rows(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                              const SellView& sv, const double* xyz_soa, int npad, double* gx,
                              double* gy, double* gz, double* lap, const std::vector<int>& rows,
                              const std::vector<int>& lens, DevBuf<int>& dev_rows);
void launch_weights_fp64(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                         const SellView& sv, const double* xyz_soa, int npad, double* gx,
                         double* gy, double* gz, double* lap, double* cond, int max_len,
                         const int* rows = nullptr, int nrows = 0, const int* only = nullptr);
// Dynamic LDS of the mixed kernel for stencils of up to max_len points.
size_t weights_mixed_lds(int max_len, int m);
// FP32 LU + FP64 refinement for the listed rows; flag[r] = 1 where it did not converge.
void launch_weights_mixed(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                          const SellView& sv, const double* xyz_soa, int npad, double* gx,
                          double* gy, double* gz, double* lap, const int* rows, int nrows,
                          int* flag, int max_len, int max_ref);

template <int DIM>
void launch_predictor(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                      const double* gz, const double* lap, const double* u, const double* v,
                      const double* w, double* uh, double* vh, double* wh, double* S_new,
                      const double* S_oo, int npad, double rho, double mu, double dt, int euler,
                      double* upk, double* uhpk, const int* bprob = nullptr,
                      const double* prm = nullptr, const double* force = nullptr);
// upk, uhpk: the velocity before and after the predictor interleaved per node
// (2 doubles per node in 2D, 4 in 3D: u, v, w, 0); the predictor packs u, v, w
// into upk for its gathers and writes uhpk beside uh, vh, wh for the
// right-hand side's gathers.
template <int DIM>
void launch_ppe_rhs(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                    const double* gz, const double* uh, const double* vh, const double* wh,
                    const double* uhpk, const double* u, const double* v, const double* w,
                    const double* p_old,
                    const uint8_t* bnd, const uint8_t* pdir, const double* normal_soa,
                    int npad, double rho, double dt, double* b, const int* bprob = nullptr,
                    const double* prm = nullptr);
template <int DIM>
void launch_corrector(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                      const double* gz, const double* p, const double* uh, const double* vh,
                      const double* wh, double* un, double* vn, double* wn, int npad, double rho,
                      double dt, const int* bprob = nullptr, const double* prm = nullptr);
template <int DIM>
void launch_boundary_fix(hipStream_t s, const SellView& sv, const int* brows, int nb,
                         const double* gx, const double* gy, const double* gz,
                         const double* normal_soa, int npad, const uint8_t* pdir,
                         const uint8_t* udir, const uint8_t* vdir, const uint8_t* wdir,
                         const double* u_old, const double* v_old, const double* w_old,
                         double* un, double* vn, double* wn);
void launch_spmv(hipStream_t s, const SellView& sv, const double* val, const double* x, double* y);
template <int DIM>
void launch_divergence(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                       const double* gz, const double* u, const double* v, const double* w,
                       const uint8_t* bnd, double* partial, int nblocks);

// Pressure operator in SELL slots: value arrays from the stencil weights.
void launch_build_operator(hipStream_t s, const SellView& sv, int dim, const double* gx,
                           const double* gy, const double* gz, const double* lap,
                           const double* normal_soa, int npad, const uint8_t* bnd,
                           const uint8_t* dir, double unsteady, double conv, double diff,
                           double* aval, double* diag);

// Largest-magnitude eigenvalue of a stencil operator (k_eig.hip); returns the
// Arnoldi restarts used.
// kNN stencils of host points into host CSR (row_ptr is filled from the
// stencil sizes); throws std::invalid_argument on bad sizes.
void knn_stencils(hipStream_t st, int nv, int dim, const double* xyz, const uint8_t* boundary,
                  const int* k_per_node, int* row_ptr, int* col);

// Reverse Cuthill-McKee order of a host CSR graph (order[old] = new); the
// number of BFS levels goes to *levels (may be null).
void rcm_order(hipStream_t st, int nv, const int* row_ptr, const int* col, int* order,
               int* levels);

int largest_eigenvalue(hipStream_t s, const SellView& sv, const double* val, int m, double tol,
                       int max_restarts, double* re, double* im);

// Krylov building blocks (k_krylov.hip).
void launch_fill(hipStream_t s, double* x, double v, int n);
void launch_copy(hipStream_t s, double* dst, const double* src, int n);

// Projection guess (k_project.hip): rings of PROJ_MAX stored pressures and
// their products with the pressure operator (vectors of ld doubles each); the
// device record holds the newest slot, the entries held and the last
// coefficients.
constexpr int PROJ_MAX = 24;
struct ProjState {
    int head;
    int count;
    double beta[PROJ_MAX];
};
size_t proj_partial_size();
// out = the combination of the newest k stored pressures (k <= PROJ_MAX, at
// most the entries held) with the smallest residual against b, over n rows
void launch_proj_guess(hipStream_t s, int k, int n, const double* hp, const double* hw, size_t ld,
                       ProjState* ps, const double* b, double* partial, double* out);
// stores p and A p = b - r as the newest entry
void launch_proj_store(hipStream_t s, int n, const double* p, const double* b, const double* r,
                       double* hp, double* hw, size_t ld, ProjState* ps);
void launch_proj_reset(hipStream_t s, ProjState* ps);

}  // namespace mp
