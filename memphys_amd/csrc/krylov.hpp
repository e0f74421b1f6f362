// Pressure-Poisson solver state (part c).  Host orchestration in k_krylov.hip.
#pragma once
#include <memory>

#include "amg.hpp"
#include "mp_internal.hpp"

namespace mp {

struct SolveStats {
    int iterations = 0, outer_cycles = 0, restarts = 0, converged = 0;
    double rel_res = 0, abs_res = 0, alpha = 0, guess_rel_res = 0;
    int gmres_switches = 0;  // GMRES runs handed to BiCGSTAB (stall or breakdown)
    int guess_rejected = 0;  // the initial guess was replaced (see PpeSolver::solve)
};

// HIP-event timing of one kernel's launches.  Krylov kernels of a system that
// has converged inside a chunk return at once; the summary counts only the
// launches that did the work: duration at least half the median launch's.
struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> ev;  // pairs
    size_t used = 0;
    std::vector<float> samples;  // milliseconds per launch
    void begin(hipStream_t s);
    void end(hipStream_t s);
    void harvest();  // after a synchronize: collect and recycle the events
    void reset() { samples.clear(); }
    // working launches, their mean duration (ms), and all launches
    void summary(long* working, double* mean_ms, long* all) const;
    void release();
};

// GMRES(m) state of one system, in device memory (m <= GM_MAX columns per
// cycle).  R holds the Givens-rotated Hessenberg matrix column by column
// (column j at R[j * GM_MAX], rows 0..j), g the rotated residual vector.
constexpr int GM_MAX = 31;
struct GmState {
    double g[GM_MAX + 1], cs[GM_MAX], sn[GM_MAX], h[GM_MAX + 1], y[GM_MAX];
    // basis vectors are stored unnormalised: v_k = sig[k] V'_k; hc[c] = h[c] sig[c]
    // is the coefficient of V'_c in the Gram-Schmidt update
    double sig[GM_MAX + 1], hc[GM_MAX];
    double R[GM_MAX * GM_MAX];
    double bb, tol2, tol_rel, rr, inv_eta;
    double rr_cycle;  // |r|^2 at the start of the current cycle
    int j;       // columns of the current cycle
    int done;    // as KScalars::done (0 running, 1 converged, 2 cap, 4 zero rhs, 5 non-finite / breakdown, 6 inactive),
                 // and 7: stalled (a restart cycle reduced |r| by less than 10%; x holds its result)
    int full;    // the cycle holds m columns: restart pending
    int stop;    // done != 0 or full: the iteration kernels return at entry
    int apply;   // x += Z y pending with napply columns
    int napply;
    int iter, maxit, cycles;
};
struct GmFlags {
    int done, full, iter, cycles;
};

// An instantiated graph of one chunk of Krylov iterations.  The key holds the
// method, the value types, the chunk's position and length and every operand
// pointer passed in; the solver's own buffers that the captured kernels also
// read (scalars, partial sums, basis, V-cycle levels) are allocated before any
// capture and released together with the graphs.
struct ChunkGraph {
    static constexpr int KEY = 12;
    uintptr_t key[KEY];
    hipGraphExec_t exec;
};

// BiCGSTAB (as Eigen 3.3.7's BiCGSTAB.h) on the right-Jacobi-scaled operator
// A D^-1, bordered with a ones row/column in the full-Neumann case, for one
// system or a batch of independent systems (the block-diagonal union of
// several clouds, each with its own scalars and stopping test).  With the
// multigrid preconditioner (one system) the operator is A itself (the scaling
// vector is all ones) and every p and s is preconditioned by one V-cycle.
struct PpeSolver {
    int nrows = 0;   // rows of all systems incl. padding (multiple of BLOCK)
    int nv = 0;      // real rows of all systems
    int n_sys = 0;   // unknowns: real rows + border unknowns
    int nseg = 1, bordered = 0, dim = 0;
    long nnz = 0;    // real stencil entries of the operator
    SellView sv{};
    // column-sorted slot order (set before setup; slot_perm null: none): the
    // FP32 operator products run on svs with their values permuted into it
    SellView svs{};
    const int* slot_perm = nullptr;
    std::vector<Seg> segs;
    DevBuf<Seg> seg_d;
    DevBuf<int> bseg_d;
    SegView sg{};
    DevBuf<double> aval, dinv;
    DevBuf<float> aval32;
    DevBuf<float> aval32s;            // aval32 in the column-sorted slot order
    const float* a32k = nullptr;      // FP32 operator of the Krylov products
    SellView svk{};                   // its slot view
    long nnz16 = 0;                   // entries of svk's slices with 16-bit columns
    int nslices16 = 0;
    DevBuf<double> xt, r, r0, p, v, s, t, tmp, e64, rhs64;
    DevBuf<double> yv, zv;            // preconditioned directions (multigrid)
    DevBuf<float> y32, z32;           // their node rows in FP32 (the values are FP32)
    DevBuf<double> jac_dinv;          // D^-1 when the operator above is unscaled
    DevBuf<float> jac_a32;            // A D^-1 in FP32 when the operator above is unscaled
    const double* diag_inv = nullptr; // D^-1 (the learned guess's features)
    const float* scaled32 = nullptr;  // A D^-1 in FP32 (the learned guess's aggregation)
    int precond = 0;                  // enum mp_precond
    int krylov = 0;                   // enum mp_krylov
    int restart_m = 30;               // GMRES columns per cycle
    DevBuf<double> gm_v, gm_w, gm_partial;  // basis (FP64 or FP32 in the same bytes), work, partials
    DevBuf<float> gm_z;               // preconditioned basis: the V-cycle's FP32 outputs
    DevBuf<GmState> gm_d;             // [nseg]
    DevBuf<GmFlags> gmf_d;            // [nseg]
    GmFlags* host_gmf = nullptr;      // pinned, 2 x nseg
    int gm_cols = 0;                  // columns allocated
    std::unique_ptr<AmgDevice> amg;
    DevBuf<double> partial;
    DevBuf<KScalars> sc;              // [nseg]
    DevBuf<SegCtl> ctl_d;             // [nseg]
    KScalars* host_sc = nullptr;      // pinned, 2 x nseg
    SegCtl* host_ctl = nullptr;       // pinned, nseg
    hipEvent_t ev[2] = {nullptr, nullptr};
    int precision = 0;
    double tol = 1e-8, inner_tol = 1e-3;
    // mixed form: each cycle's inner solve ends at target_margin * tol (0.5,
    // or 0.25 under the projected guess, whose next guess starts below the
    // tolerance more often when this solve ends further below it)
    double target_margin = 0.5;
    int maxit = 1000, check_every = 8;
    // GMRES iterations of the last run (0: none yet): the next run's first
    // chunk stops one short of it, and the chunks after it are of two
    // iterations, so a solve that converges as fast as the last one leaves
    // few queued iterations to return at once
    int gm_hint = 0;
    long fp64_fallbacks = 0;
    KernelTimer spmv_timer;
    KernelTimer smooth_timer;         // the multigrid's level-0 post-smoothing product
    bool use_graphs = true;           // replay each chunk of iterations from a hipGraph
    std::vector<ChunkGraph> graphs;
    long total_iterations = 0, total_solves = 0;
    std::vector<SolveStats> seg_stats;  // per system, of the last solve

    ~PpeSolver();
    // Builds the operator from weights in SELL slots; view.nv = nrows.
    void setup(hipStream_t s, const SellView& view, const std::vector<Seg>& systems, int dim,
               const double* gx, const double* gy, const double* gz, const double* lap,
               const double* normal_soa, const uint8_t* bnd_dev, const uint8_t* dir_dev,
               double unsteady, double conv, double diff, int precond = 0);
    // b_dev, x_dev: device rows (guess in x_dev, solution out).  Returns the
    // statistics of a single system, or the worst over a batch (per system in
    // seg_stats).
    SolveStats solve(hipStream_t s, const double* b_dev, double* x_dev);
    double spmv_bytes(int fp32) const;
    // algorithmic bytes of the level-0 post-smoothing product (BF16 values)
    double smooth_bytes() const;
    // r = b - A x (FP64, unscaled operator incl. border) into `tmp`; returns tmp.p
    double* residual_of(hipStream_t s, const double* b_dev, const double* x_dev);
    // b - A x for the x the last solve() returned (FP64, in `tmp`, valid until
    // the next solve or residual_of)
    const double* final_residual() const { return tmp.p; }
    // |b - A x|^2 of system 0 from the last residual_of (waits for the stream)
    double last_residual_norm2(hipStream_t s);
    // Second candidate for the next solve's guess (single system): solve()
    // starts from it instead when the guess's FP64 residual is not finite or
    // exceeds FALLBACK_FACTOR times the candidate's (fallback_rr = its
    // |b - A x|^2).  Cleared by solve().  The factor is not 1: the learned
    // guess typically leaves a larger residual than its extrapolation (its
    // remaining error is high-frequency, which the Krylov solve removes in a
    // few iterations) and still converges in fewer iterations; the bound
    // catches a broken network, not a worse-looking residual.
    static constexpr double FALLBACK_FACTOR = 10.0;
    const double* fallback_x = nullptr;
    double fallback_rr = 0.0;
    // Drops the captured chunk graphs (their kernel arguments point into the
    // operator and preconditioner buffers).
    void release_graphs();

  private:
    template <typename TM>
    const SellView& view_of() const {  // the slot view of operator values of type TM
        return sizeof(TM) == 8 ? sv : svk;
    }
    template <typename TM, typename T>
    void run_bicgstab(hipStream_t s, const TM* aval_t, const T* dinv_t, const T* rhs, T* x,
                      T* rr, T* rr0, T* pp, T* vv, T* ss, T* tt, int rel_to_rhs,
                      std::vector<int>& done, std::vector<int>& iters, std::vector<int>& restarts);
    // GMRES(restart_m) on operator values aval_t (FP64 x and right-hand
    // side, basis vectors of type TV), right-preconditioned by the V-cycle
    // when there is one; same controls and outcome codes as run_bicgstab
    // (restarts = GMRES cycles after the first).  x_zero: x is zero on entry
    // (the first residual is rhs itself).
    template <typename TM, typename TV>
    void run_gmres(hipStream_t s, const TM* aval_t, const double* dinv_t, const double* rhs,
                   double* x, int rel_to_rhs, std::vector<int>& done, std::vector<int>& iters,
                   std::vector<int>& restarts, bool x_zero = false);
    template <typename TM, typename TV>
    void gmres_iteration(hipStream_t s, int k, int m, const TM* aval_t, const double* dinv_t);
    void gmres_alloc(hipStream_t s, int m);
    void residual64(hipStream_t s, const double* b_dev, std::vector<double>& rr,
                    std::vector<double>& bb);
    const double* precondition(hipStream_t s, const double* in, double* out, float* out32);
    template <typename TM, typename T>
    void launch_chunk(hipStream_t s, int chunk, const TM* a, const T* di, T* x, T* rr, T* rr0,
                      T* pp, T* vv, T* ss, T* tt);
    template <typename TM, typename T>
    void run_chunk(hipStream_t s, int chunk, const TM* a, const T* di, T* x, T* rr, T* rr0,
                   T* pp, T* vv, T* ss, T* tt);
    // Iterations j .. j+c-1 of a GMRES cycle: replayed from a cached graph
    // when graphs are on, else launched.
    template <typename TM, typename TV>
    void run_gmres_chunk(hipStream_t s, int j, int c, int m, const TM* a, const double* di);
    // Replays the graph cached under key, or captures launch() on the stream,
    // instantiates it, caches it and replays it.
    template <typename F>
    void replay_or_capture(hipStream_t s, const uintptr_t (&key)[ChunkGraph::KEY], F&& launch);
    void zero_system(hipStream_t s, double* x, int q);
};

}  // namespace mp
