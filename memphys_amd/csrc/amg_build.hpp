// Smoothed-aggregation algebraic multigrid: host set-up of the hierarchy that
// preconditions the pressure-Poisson BiCGSTAB.  Plain C++ (no HIP), so the
// same source builds into the host library for CPU checks.
//
// Level l -> l+1: strength graph = the k largest |a_ij| (j != i) of each row,
// symmetrised; greedy aggregation in row order (a node whose strong
// neighbours are all free seeds an aggregate with them; leftovers join the
// aggregate of their first aggregated neighbour, else stay alone); tentative
// prolongator = aggregate indicator (the constant near-null space of the
// Neumann Laplacian); smoothed P = (I - omega D^-1 A) P_tent, optionally
// truncated (entries below trunc x the row's largest |p| dropped, the rest
// rescaled to keep the row sum, so constants are still reproduced); R = P^T;
// A_{l+1} = R A_l P (Galerkin), optionally sparsified (off-diagonal entries
// below drop x sqrt|a_ii a_jj| moved onto the diagonal, which keeps row sums
// and so the constant null vector).  The coarsest level of a full-Neumann
// (bordered) system is solved with the dense inverse of its matrix bordered by
// a ones row and column (the constant mode is singular there); a system with
// Dirichlet rows is nonsingular and uses the plain dense inverse.
#pragma once
#include <vector>

namespace mp_amg {

struct Csr {
    int nrows = 0, ncols = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
    long nnz() const { return ptr.empty() ? 0 : ptr[nrows]; }
};

struct Options {
    int strength_k = 6;      // strongest off-diagonal couplings kept per row
    double omega_p = 0.66;   // prolongation smoothing weight
    double trunc = 0.15;     // prolongator truncation (relative to the row's max)
    double drop = 0.01;      // coarse-operator sparsification threshold
    int coarse_max = 700;    // a level with at most this many rows is solved densely
    int max_levels = 12;
    int threads = 0;         // 0: hardware concurrency
};

struct Hierarchy {
    // A[0] is left empty (the caller owns the finest operator); A[l] for
    // l >= 1 are the Galerkin operators.  P[l], R[l] map between l and l+1.
    std::vector<Csr> A, P, R;
    std::vector<double> coarse_inv;  // nc x nc, row-major
    int nc = 0;
    int levels() const { return (int)A.size(); }
};

void aggregate(const Csr& A, int k, std::vector<int>& agg, int& n_agg);
Csr smoothed_prolongator(const Csr& A, const std::vector<int>& agg, int n_agg, double omega,
                         double trunc = 0.0);
// Moves off-diagonal entries below drop * sqrt(|a_ii a_jj|) onto the diagonal.
Csr sparsify(const Csr& A, double drop);
Csr transpose(const Csr& A);
Csr multiply(const Csr& A, const Csr& B, int threads);
// Top-left n x n block of the inverse of [[A, 1], [1^T, 0]] (A dense n x n).
std::vector<double> bordered_inverse(const Csr& A);
// Inverse of A (dense n x n, row-major); throws when A is singular.
std::vector<double> dense_inverse(const Csr& A);
// bordered: the coarsest solve is bordered_inverse (full-Neumann system),
// else dense_inverse.
Hierarchy build(const Csr& A0, const Options& opt, bool bordered = true);

}  // namespace mp_amg
