// Internal declarations shared by the HIP translation units of libmemphys_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/memphys_hip.h"

namespace mp {

constexpr int SLICE = 64;       // rows per SELL slice (one wavefront)
constexpr int SIGMA = 4096;     // rows per length-sorting window
constexpr int BLOCK = 256;      // threads per block of the row kernels
constexpr int MAX_STENCIL = 96; // largest stencil (k) the weight kernel accepts
constexpr int MAX_POLY = 35;    // largest monomial count (degree 4 in 3D)

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

// SELL-64 view of the stencil set in device order.
struct SellView {
    int nv;            // real rows
    int nslices;
    const int* sptr;   // [nslices+1] first slot of each slice
    const int* col;    // [slots] device column ids
    const int* len;    // [nslices*64] row lengths (0 on ghost rows)
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
void launch_weights_fp64(hipStream_t s, int dim, int phs, int m, const int* exps_dev,
                         const SellView& sv, const double* xyz_soa, int npad, double* gx,
                         double* gy, double* gz, double* lap, double* cond, int max_len,
                         const int* rows = nullptr, int nrows = 0);
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
                      const int* bprob = nullptr, const double* prm = nullptr);
template <int DIM>
void launch_ppe_rhs(hipStream_t s, const SellView& sv, const double* gx, const double* gy,
                    const double* gz, const double* uh, const double* vh, const double* wh,
                    const double* u, const double* v, const double* w, const double* p_old,
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
int largest_eigenvalue(hipStream_t s, const SellView& sv, const double* val, int m, double tol,
                       int max_restarts, double* re, double* im);

// Krylov building blocks (k_krylov.hip).
void launch_fill(hipStream_t s, double* x, double v, int n);

}  // namespace mp
