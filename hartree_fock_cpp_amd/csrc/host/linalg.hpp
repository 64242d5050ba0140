// linalg.hpp -- dense FP64 linear algebra for the host side of the SCF.
//
// Row-major K x K matrices in std::vector<double>, the layout of the
// reference (include/indexing.hpp:4-7).  No external libraries: the
// symmetric eigensolver is Householder tridiagonalisation followed by the
// implicit symmetric QR iteration with Wilkinson shifts (Golub & Van Loan,
// Matrix Computations, 4th ed., Algorithms 8.3.1 and 8.3.3); the DIIS system
// is solved by LU with full pivoting (GvL Algorithm 3.4.2).
#pragma once
#include <vector>

namespace hfxh {

using Mat = std::vector<double>;

// C = A * B (n x n); ta/tb transpose A/B.
void gemm(int n, const double* A, bool ta, const double* B, bool tb, double* C);

// Eigen-decomposition of a symmetric n x n matrix A (row-major):
// w ascending, V row-major with eigenvector k in column k.
void sym_eig(int n, const double* A, double* w, double* V);

// All eigenvalues (ascending, in w) and the eigenvectors of the m lowest
// (columns 0..m-1 of V; the other columns are zero): eigenvalues by the QR
// iteration without vectors, eigenvectors by inverse iteration on the
// tridiagonal matrix with Gram-Schmidt inside clusters (LAPACK dstein's
// scheme), transformed back through the Householder reflectors.  About
// 4/3 n^3 + 4 m n^2 flops against ~8 n^3 for sym_eig.
void sym_eig_lowest(int n, const double* A, int m, double* w, double* V);

// Solves A x = b (n x n, row-major) by LU with full pivoting.  Singular
// pivots are treated as zero components of x (the minimum-effort solution
// on the pivot's subspace).
void solve_fullpiv(int n, const double* A, const double* b, double* x);

}  // namespace hfxh
