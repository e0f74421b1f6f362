// Small dense eigenvalue routines for the Arnoldi iteration of k_eig.hip
// (plain C++, also built into the host library for CPU checks).
#pragma once
#include <complex>
#include <vector>

namespace mp_eig {

using cplx = std::complex<double>;

// Eigenvalues of an upper Hessenberg matrix (row-major m x m) by the shifted
// QR iteration in complex arithmetic with Wilkinson shifts and deflation.
std::vector<cplx> hessenberg_eigenvalues(const std::vector<double>& H, int m);

// Unit eigenvector of H for the eigenvalue lam by two steps of inverse
// iteration (complex Gaussian elimination with partial pivoting).
std::vector<cplx> hessenberg_eigenvector(const std::vector<double>& H, int m, cplx lam);

}  // namespace mp_eig
