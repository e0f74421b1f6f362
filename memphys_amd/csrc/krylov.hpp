// Pressure-Poisson solver state (part c).  Host orchestration in k_krylov.hip.
#pragma once
#include <memory>

#include "amg.hpp"
#include "mp_internal.hpp"

namespace mp {

struct SolveStats {
    int iterations = 0, outer_cycles = 0, restarts = 0, converged = 0;
    double rel_res = 0, abs_res = 0, alpha = 0, guess_rel_res = 0;
};

struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> ev;  // pairs
    size_t used = 0;
    long launches = 0;
    double total_ms = 0.0;
    void begin(hipStream_t s);
    void end(hipStream_t s);
    void harvest();  // after a synchronize: accumulate and recycle
    void release();
};

// An instantiated graph of one chunk of BiCGSTAB iterations for one set of
// operand pointers.
struct ChunkGraph {
    const void* key[9];
    int chunk;
    hipGraphExec_t exec;
};

// BiCGSTAB (as Eigen 3.3.7's BiCGSTAB.h) on the right-Jacobi-scaled operator
// A D^-1, bordered with a ones row/column in the full-Neumann case.  With the
// multigrid preconditioner the operator is A itself (the scaling vector is
// all ones) and every p and s is preconditioned by one V-cycle.
struct PpeSolver {
    int nv = 0, npad = 0, n = 0, bordered = 0, dim = 0;
    long nnz = 0;  // real stencil entries of the operator
    SellView sv{};
    DevBuf<double> aval, dinv;
    DevBuf<float> aval32;
    DevBuf<double> xt, r, r0, p, v, s, t, tmp, e64, rhs64;
    DevBuf<double> yv, zv;            // preconditioned directions (multigrid)
    DevBuf<double> jac_dinv;          // D^-1 when the operator above is unscaled
    DevBuf<float> jac_a32;            // A D^-1 in FP32 when the operator above is unscaled
    const double* diag_inv = nullptr; // D^-1 (the learned guess's features)
    const float* scaled32 = nullptr;  // A D^-1 in FP32 (the learned guess's aggregation)
    int precond = 0;                  // enum mp_precond
    std::unique_ptr<AmgDevice> amg;
    DevBuf<double> partial;
    DevBuf<KScalars> sc;
    KScalars* host_sc = nullptr;  // pinned, 2 slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    int precision = 0;
    double tol = 1e-8, inner_tol = 1e-3;
    int maxit = 1000, check_every = 8, inner_maxit = 1000;
    long fp64_fallbacks = 0;
    KernelTimer spmv_timer;
    bool use_graphs = true;           // replay each chunk of iterations from a hipGraph
    std::vector<ChunkGraph> graphs;
    long total_iterations = 0, total_solves = 0;

    ~PpeSolver();
    // Builds the operator from weights in SELL slots.
    void setup(hipStream_t s, const SellView& view, int dim, int npad, const double* gx,
               const double* gy, const double* gz, const double* lap, const double* normal_soa,
               const uint8_t* bnd_dev, const uint8_t* dir_dev, bool bordered, double unsteady,
               double conv, double diff, int precond = 0);
    // b_dev, x_dev: device order, length n (guess in x_dev, solution out).
    SolveStats solve(hipStream_t s, const double* b_dev, double* x_dev);
    double spmv_bytes(int fp32) const;
    // r = b - A x (FP64, unscaled operator incl. border) into `tmp`; returns tmp.p
    double* residual_of(hipStream_t s, const double* b_dev, const double* x_dev);

  private:
    template <typename TM, typename T>
    int run_bicgstab(hipStream_t s, const TM* aval_t, const T* dinv_t, const T* rhs, T* x, T* rr,
                     T* rr0, T* pp, T* vv, T* ss, T* tt, double tol_rel, int rel_to_rhs, int maxit,
                     int* restarts, int* done_out);
    void residual64(hipStream_t s, const double* b_dev, double* out_rr, double* out_bb);
    const double* precondition(hipStream_t s, const double* in, double* out);
    template <typename TM, typename T>
    void launch_chunk(hipStream_t s, int chunk, const TM* a, const T* di, T* x, T* rr, T* rr0,
                      T* pp, T* vv, T* ss, T* tt);
    template <typename TM, typename T>
    void run_chunk(hipStream_t s, int chunk, const TM* a, const T* di, T* x, T* rr, T* rr0,
                   T* pp, T* vv, T* ss, T* tt);
    void release_graphs();
};

}  // namespace mp
