/* memphys_hip.h — C ABI of the MI355X (gfx950) meshless fractional-step path.
 *
 * Plain pointers and sizes only.  Host pointers are read or written before the
 * call returns; all device state lives in the context.  Every entry point
 * returns 0 on success and a non-zero mp_status otherwise; mp_last_error()
 * then describes the failure.  One context per host thread; calls on one
 * context are not re-entrant.  The C++ class layer (memphys_amd/host) turns a
 * non-zero status into the reference's printf + throw bad_exception().
 *
 * Indices are 32-bit (the reference's CSR uses int, class.hpp:145-146).
 * Node order at this boundary is the caller's (the RCM order CLOUD produces);
 * the library keeps its own device order internally and translates.
 */
#ifndef MEMPHYS_HIP_H
#define MEMPHYS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mp_ctx mp_ctx;

enum mp_status {
    MP_OK = 0,
    MP_ERR_ARG = 1,      /* invalid argument or size */
    MP_ERR_HIP = 2,      /* HIP runtime error */
    MP_ERR_STATE = 3,    /* call out of order (e.g. step before create) */
    MP_ERR_NOCONV = 4    /* Krylov solve stopped above tolerance */
};

/* Fields of mp_fields_set/get.  MP_SU..MP_SW: the AB2 source history (the
 * explicit right-hand side S of the last step, S_old of fractional_step_1.cpp
 * :120-134), which the next AB2 step combines with its own; with the
 * velocities, the pressure and the step counter it makes a restart exact
 * (no counterpart in the reference, whose tecplot restart starts AB2 anew). */
enum mp_field { MP_U = 0, MP_V = 1, MP_W = 2, MP_P = 3, MP_SU = 4, MP_SV = 5, MP_SW = 6 };
enum mp_op { MP_GX = 0, MP_GY = 1, MP_GZ = 2, MP_LAP = 3 };

/* PPE arithmetic: plain FP64 BiCGSTAB, or FP32 inner BiCGSTAB inside FP64
 * defect correction (outer residual b - A x in FP64). */
enum mp_precision { MP_PREC_FP64 = 0, MP_PREC_MIXED = 1 };

/* Initial guess for each pressure solve: the previous pressure (reference
 * behaviour, solveWithGuess(rhs, p_old)), the learned message-passing guess
 * (a classical guess, mp_guess_base, plus the network's correction), the
 * extrapolation alone: the polynomial through the pressures of the last 2, 3
 * or 4 steps (mp_guess_order; equal time steps), evaluated at the next step,
 * or the projection: the combination of the last K solutions (mp_guess_window,
 * default 20) whose residual against the new right-hand side is smallest
 * (Fischer's successive-right-hand-side projection; A p of each solution is
 * kept from its solve's final FP64 residual).  The extrapolating guesses
 * fall back to the previous pressure until two pressures are in the history,
 * and use the highest order the history holds until it is full; the
 * projection uses the solutions held so far.  The learned and projected
 * guesses are for single problems; a batch may extrapolate (every problem's
 * rows, border unknowns included).  No counterpart in the reference beyond
 * MP_GUESS_PREVIOUS. */
enum mp_guess {
    MP_GUESS_PREVIOUS = 0,
    MP_GUESS_LEARNED = 1,
    MP_GUESS_EXTRAPOLATE = 2,
    MP_GUESS_PROJECT = 3
};

/* PPE preconditioner: Jacobi (right diagonal scaling, the plain path), or a
 * smoothed-aggregation algebraic multigrid V(1,1)-cycle in FP32 (set up on the
 * host from the assembled operator, applied on the device). */
enum mp_precond { MP_PRECOND_JACOBI = 0, MP_PRECOND_AMG = 1 };

/* PPE Krylov method: BiCGSTAB as Eigen's (the reference's eigen_ilu_bicgstab,
 * solver.cpp:337-350), or restarted GMRES (the reference's hypre_ilu_gmres,
 * solver.cpp:41-53 and 140-144, with gmres_kdim as the restart length),
 * right-preconditioned, classical Gram-Schmidt, Givens rotations. */
enum mp_krylov { MP_KRYLOV_BICGSTAB = 0, MP_KRYLOV_GMRES = 1 };

typedef struct {
    double rho, mu, dt;
    int temporal_order;       /* 1 = Euler, 2 = AB2 after the first step */
    double solver_tolerance;  /* relative residual |b - A p| / |b| */
    int n_iter;               /* Krylov iteration cap per solve */
    int precision;            /* enum mp_precision */
    double inner_tolerance;   /* MIXED: inner FP32 relative reduction per cycle */
    int check_every;          /* iterations launched between convergence polls (0 = default) */
    int preconditioner;       /* enum mp_precond */
    int krylov;               /* enum mp_krylov, 0 = BiCGSTAB */
    int restart;              /* GMRES columns per cycle, 1..31 (0 = 30) */
} mp_fs_params;

typedef struct {
    int iterations;       /* Krylov iterations of the last solve (all cycles) */
    int outer_cycles;     /* FP64 defect-correction cycles (1 for FP64) */
    int restarts;         /* BiCGSTAB restarts */
    int converged;        /* 1 when rel_res <= tolerance */
    double rel_res;       /* |A p - b| / |b| of the final pressure (fresh FP64 product) */
    double abs_res;       /* |A p - b| */
    double alpha;         /* regularisation unknown p[nv] (full-Neumann only, else 0) */
    double guess_rel_res; /* |A p0 - b| / |b| of the initial guess */
    long total_iterations;/* sum over all solves since context creation */
    long total_solves;
    int gmres_switches;   /* GMRES cycles that stalled (|r| reduced by < 10%) or broke down,
                             after which the system continued with BiCGSTAB */
    int guess_rejected;   /* 1 when the initial guess was replaced: a non-finite guess of a
                             finite system starts from zero, a learned guess whose residual is
                             not finite or exceeds 10x its linear extrapolation's starts from
                             the extrapolation */
    int guess_skipped;    /* 1 when the learned guess was not evaluated: after a rejected
                             guess the network is skipped for 1, 2, 4 ... 64 steps (the
                             extrapolation alone); an accepted guess resets the back-off */
} mp_step_stats;

/* Context ------------------------------------------------------------------ */
int mp_ctx_create(int device, mp_ctx** out);
int mp_ctx_destroy(mp_ctx* ctx);
const char* mp_last_error(const mp_ctx* ctx);
int mp_sync(mp_ctx* ctx);

/* Point cloud and stencils (replaces the CLOUD data handed to the weight loop,
 * cloud.cpp:207-279; POINTS xyz/normal/boundary_flag, class.hpp:99-124).
 * xyz and normal: nv*dim AoS; boundary: nv bytes; CSR row_ptr[nv+1], col[...];
 * entry 0 of every row must be the node itself; rows of 1..512 entries. */
int mp_cloud_upload(mp_ctx* ctx, int nv, int dim, const double* xyz, const uint8_t* boundary,
                    const double* normal, const int* row_ptr, const int* col_idx);

/* Batch of independent clouds (no reference counterpart: the reference runs
 * one cloud per process).  nprob clouds given back to back in caller order:
 * problem q owns caller nodes [off_q, off_q + nv_per_problem[q]) with off_q the
 * sum of the earlier sizes; row_ptr/col_idx index the concatenation and a
 * stencil may only reference nodes of its own problem.  Every later call
 * (weights, mp_fs_create, mp_fs_step, fields) then serves all problems at once:
 * one kernel launch per phase covers every problem, and each pressure system
 * keeps its own Krylov scalars and stopping test.  Node fields are the
 * concatenation of the problems' fields; a pressure array is the nodes
 * followed by the border unknowns of the full-Neumann problems in problem
 * order.  mp_cloud_upload is the one-problem case. */
int mp_batch_upload(mp_ctx* ctx, int nprob, const int* nv_per_problem, int dim, const double* xyz,
                    const uint8_t* boundary, const double* normal, const int* row_ptr,
                    const int* col_idx);

/* Batched PHS + polynomial stencil weights (replaces CLOUD::calc_grad_laplace_coeffs,
 * cloud.cpp:207-279, and calc_PHS_RBF_grad_laplace_single_vert,
 * coefficient_computations.cpp:294-328).  exps: m*dim monomial exponents.
 * precision_mode: MP_PREC_FP64, or MP_PREC_MIXED (FP32 LU + FP64 refinement, FP64
 * fallback per stencil).  Output arrays (CSR layout of the upload, may be NULL
 * to keep the weights device-resident): gx, gy, gz (3D), lap; cond_out[nv]
 * (may be NULL) receives a 1-norm condition estimate of each stencil matrix.
 * poly_deg: any degree with m <= 136 monomials (2D up to degree 15, the reference's
 * range; 3D up to degree 7); systems of M = k + m too large for LDS are solved from a global
 * work area, more slowly, and the mixed mode solves them in FP64. */
int mp_weights_assemble(mp_ctx* ctx, int phs_deg, int poly_deg, const int* exps,
                        int precision_mode, double* gx, double* gy, double* gz, double* lap,
                        double* cond_out);

/* Upload weights computed elsewhere (CSR layout of the upload). */
int mp_weights_set(mp_ctx* ctx, const double* gx, const double* gy, const double* gz,
                   const double* lap);

/* y = Op x for one stencil operator (replaces points.*_matrix_EIGEN * x). */
int mp_operator_apply(mp_ctx* ctx, int op, const double* x, double* y);

/* Stencil selection on the GPU (replaces CLOUD::calc_cloud_points_fast,
 * cloud.cpp:570-648): an interior node gets its k nearest nodes (itself
 * first), a boundary node itself and its k-1 nearest interior nodes; each
 * row is ordered by (squared distance, index), exact ties going to the lower
 * index.  Host arrays: xyz nv*dim (AoS), boundary nv, k_per_node nv (1..512).
 * Output CSR in host memory: row_ptr nv+1 (filled from k_per_node) and col
 * (row_ptr[nv] entries).  Needs no uploaded cloud; leaves the context's
 * cloud untouched. */
int mp_knn_stencils(mp_ctx* ctx, int nv, int dim, const double* xyz, const uint8_t* boundary,
                    const int* k_per_node, int* row_ptr, int* col);

/* Reverse Cuthill-McKee order on the GPU (replaces cuthill_mckee_ordering +
 * reverse_cuthill_mckee_ordering, general_functions.cpp:807-879, as called by
 * CLOUD::re_order_points_reverse_cuthill_mckee, cloud.cpp:315-331): start at
 * the lowest-index node of minimum degree (row length), visit each node's
 * stencil neighbours, append every newly visited set sorted by (degree,
 * index), reverse.  Nodes the start cannot reach restart the search at the
 * lowest-index unvisited node of minimum degree.  Host CSR in (row_ptr nv+1,
 * col), host order out: order[old] = new. */
int mp_rcm_order(mp_ctx* ctx, int nv, const int* row_ptr, const int* col, int* order);

/* Largest-magnitude eigenvalue of one stencil operator (replaces
 * calc_largest_magnitude_eigenvalue, general_functions.cpp:4-24, as used by
 * PARAMETERS::calc_dt, parameters.cpp:269-296): restarted Arnoldi on the GPU
 * until the Ritz residual is below tol relative; im >= 0 (a conjugate pair is
 * reported once). */
int mp_operator_eigenvalue(mp_ctx* ctx, int op, double tol, int max_restarts, double* re,
                           double* im);

/* Implicit operator and Krylov solve (replaces SOLVER::init, solver.cpp:14-79 with
 * calc_coeff_matrix 165-214, and SOLVER::general_solve 283-363).
 * Rows: interior diff*L + conv*(Gx+Gy+Gz) + unsteady*I; boundary Dirichlet
 * identity; boundary Neumann n.grad.  With no Dirichlet boundary row the system
 * is bordered to nv+1 (ones column and row, zero corner). */
int mp_solver_init(mp_ctx* ctx, const uint8_t* dirichlet, double unsteady, double conv,
                   double diff, double tolerance, int n_iter, int precision);
/* Krylov method of the solver made by mp_solver_init (SOLVER::init's Hypre
 * GMRES branch, solver.cpp:41-53 and 140-144, restart length gmres_kdim):
 * enum mp_krylov and the GMRES restart length 1..31 (0 = 30).  Call after
 * mp_solver_init; the default is BiCGSTAB. */
int mp_solver_krylov(mp_ctx* ctx, int krylov, int restart);
int mp_solver_system_size(mp_ctx* ctx, int* n);
/* field_new = solve(rhs) from the guess field_old; field_old[nv] = field_new[nv]
 * when bordered.  Arrays of system size. */
int mp_solver_solve(mp_ctx* ctx, const double* rhs, double* field_new, double* field_old);

/* Fractional step (replaces FRACTIONAL_STEP_1, fractional_step_1.cpp:22-222,
 * generalised to 3D).  Builds the pressure operator as mp_solver_init with
 * p_dir and (0, 0, 1).  w_dir is ignored in 2D. */
int mp_fs_create(mp_ctx* ctx, const mp_fs_params* params, const uint8_t* u_dir,
                 const uint8_t* v_dir, const uint8_t* w_dir, const uint8_t* p_dir);
int mp_fields_set(mp_ctx* ctx, int which, const double* values);
int mp_fields_get(mp_ctx* ctx, int which, double* values);
/* Body force per unit volume (nv entries each, caller order) added to the
 * explicit source of every following step: the overload of
 * FRACTIONAL_STEP_1::single_timestep_2d with body_force_x/y (class.hpp:253,
 * fractional_step_1.cpp:65-76 and 104-118), generalised to 3D (fz).  All
 * NULL: no body force (the plain overload). */
int mp_fs_body_force(mp_ctx* ctx, const double* fx, const double* fy, const double* fz);
/* n_steps of single_timestep followed by old <- new; it0 is the first step index
 * (Euler at it == 0).  Fields stay on the device between calls. */
int mp_fs_step(mp_ctx* ctx, int n_steps, int it0);
int mp_stats_get(mp_ctx* ctx, mp_step_stats* stats);
int mp_guess_mode(mp_ctx* ctx, int mode);
/* Order of the extrapolated guess (2: 2 p1 - p2, 3: 3 (p1 - p2) + p3, 4:
 * 4 (p1 + p3) - 6 p2 - p4, p1 the latest pressure); default 4.  The higher
 * orders amplify the error the history carries from the solves (by 2.2, 4.4
 * and 8.3 in norm): with loose tolerances a lower order can start closer.
 * No counterpart in the reference (its guess is the previous pressure). */
int mp_guess_order(mp_ctx* ctx, int order);
/* Solutions the projected guess combines (1..24, default 20). */
int mp_guess_window(mp_ctx* ctx, int k);
/* The classical guess the learned network corrects: MP_GUESS_PROJECT (the
 * default) or MP_GUESS_EXTRAPOLATE.  It is also the second candidate the
 * solve falls back to. */
int mp_guess_base(mp_ctx* ctx, int base);

/* Learned guess: a 3-layer message-passing network over the stencil graph
 * (width 64, BF16 weights, FP32 accumulation).  The parameters are a packed
 * FP32 vector of mp_guess_param_count() entries.
 * mp_guess_train: while enabled, every pressure solve of mp_fs_step trains the
 * network on (features before the solve -> converged pressure) with one Adam
 * step; seed != 0 re-initialises the parameters first.  last_loss receives the
 * mean squared error of the last training step. */
int mp_guess_param_count(void);
int mp_guess_train(mp_ctx* ctx, int enable, double learning_rate, unsigned long long seed);
int mp_guess_params(mp_ctx* ctx, int set, float* params);
int mp_guess_last_loss(mp_ctx* ctx, double* loss);
/* One evaluation on caller-order inputs (needs mp_fs_create): guess (nv
 * entries) from p1 = p(t-1), p2 = p(t-2) and the right-hand side b (system
 * size), with the network's correction on top of 2 p1 - p2 (the border
 * unknown taken from p1 in the residual feature); with p_true !=
 * NULL also the loss and its gradient (packed like the parameters).  No
 * parameter update. */
int mp_guess_eval(mp_ctx* ctx, const double* p1, const double* p2, const double* b,
                  const double* p_true, double* guess, float* grad, double* loss);

/* Continuity residual Gx u + Gy v (+ Gz w) of the current fields
 * (postprocessing_functions.cpp:192-195, 248-251):
 * out = {max_int, mean_int, max_bnd, mean_bnd, max_all, mean_all}. */
int mp_divergence(mp_ctx* ctx, double* out);

/* mp_divergence without allocating (no reference counterpart; for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops).  out_dev: 6 doubles in DEVICE memory, written on the context's
 * stream in the same order and with the same values as mp_divergence; the
 * call does not wait for the stream.  The per-block partial sums live in a
 * fixed device array of the library, one per GPU and shared by all contexts on
 * it: clouds of more than MP_NOALLOC_MAX_ROWS device rows (nodes padded to
 * whole blocks of 256 per problem) are refused with MP_ERR_ARG, and calls that
 * overlap on one GPU (other threads or streams) race on that array. */
#define MP_NOALLOC_MAX_ROWS 4194304
int mp_divergence_noalloc(mp_ctx* ctx, double* out_dev);

/* Per-problem rho, mu, dt of a batch (after mp_fs_create; arrays of nprob);
 * replaces the scalars of mp_fs_params for the time steps. */
int mp_batch_params(mp_ctx* ctx, const double* rho, const double* mu, const double* dt);
/* Statistics of one problem's last pressure solve (total_iterations: since
 * mp_fs_create). */
int mp_batch_stats(mp_ctx* ctx, int problem, mp_step_stats* stats);

/* Autotuner (no reference counterpart): after at least one mp_fs_step,
 * (1) with the multigrid preconditioner, picks the layout of every hierarchy
 * matrix (SELL-64 or CSR with 4..64 lanes per row) by timing its product; the
 * layouts sum rows in different orders, so later V-cycles differ from the
 * set-up's rule-of-thumb layouts at the rounding level;
 * (2) times the last step's pressure solve for convergence-poll intervals of
 * 4, 8, 16, 32 and 64 iterations (each from the same starting pressure; the
 * results do not depend on the interval) and keeps the fastest.  chosen
 * receives it, ms_per_candidate (may be NULL) the 5 timings.  Fields and the
 * pressure history are left unchanged. */
int mp_autotune(mp_ctx* ctx, int* chosen, double* ms_per_candidate);

/* Multigrid hierarchy of the pressure operator (MP_PRECOND_AMG): number of
 * levels (<= max_levels), rows and stored entries per level, host set-up time. */
int mp_precond_info(mp_ctx* ctx, int max_levels, int* levels, int* rows, long* nnz,
                    double* setup_seconds);

/* One application of the pressure solve's multigrid preconditioner (the FP32
 * V(1,1)-cycle of MP_PRECOND_AMG): out = M in, both of the system's size n
 * (nv node rows in caller order, then the border unknowns, which pass
 * through).  cycle 0:
 * the cycle the solver runs (with three or more levels, the coarse levels
 * from composed matrices); 1: the level-by-level cycle, the same linear map
 * with each factor rounded on its own.  No counterpart in the reference (its
 * preconditioners live inside Eigen's solvers): for callers that
 * precondition a solver of their own, and for checks of the cycle. */
int mp_precond_apply(mp_ctx* ctx, int cycle, int n, const double* in, double* out);

/* Krylov iterations between two convergence polls are replayed from a
 * hipGraph captured on first use (enable = 1, the default) or launched one by
 * one (enable = 0).  Results are identical; kernel timing (below) always
 * launches one by one. */
int mp_graph_mode(mp_ctx* ctx, int enable);

/* Per-kernel timing with HIP events on the library's stream.
 * which: 0 = PPE SpMV, 1 = predictor, 2 = pressure RHS, 3 = corrector,
 *        4 = stencil weights, 5 = the multigrid's level-0 post-smoothing
 *        product (with MP_PRECOND_AMG).  Reports the launches that did the work and
 * their mean milliseconds: Krylov launches of a system that converged inside
 * a chunk of iterations return at once and are left out (a launch counts
 * when it takes at least a quarter of the longest launch's time).  all_launches (may
 * be NULL) receives every timed launch.  bytes_per_launch: the kernel's
 * algorithmic HBM bytes per working launch. */
int mp_kernel_timing(mp_ctx* ctx, int enable);
int mp_kernel_timing_get(mp_ctx* ctx, int which, long* launches, double* mean_ms,
                         double* bytes_per_launch, long* all_launches);

#ifdef __cplusplus
}
#endif
#endif /* MEMPHYS_HIP_H */
