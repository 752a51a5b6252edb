// eigen.hpp -- block vector kernels (tall-skinny dot products and linear
// combinations over columns given by pointer) and the LOBPCG eigensolver
// for the Jacobian (eigen.hip).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "handles.hpp"

namespace nosh {

// Columns per block operation.
constexpr int kMaxCols = 48;

// Scratch of the block operations (grown on demand).
struct BlockWork {
  DBuf<double> partials, out, coef;
  std::vector<double> host;
};

// out (host, ma x mb, row-major) = A^T B over L rows; columns by pointer.
// Deterministic: fixed per-block partial sums, then a fixed-order sum.
void block_dots(const double* const* A, int ma, const double* const* B, int mb, int64_t L,
                BlockWork& w, double* out, hipStream_t s);

// Y_j = beta Y_j + sum_i C[i k + j] S_i for j < k (C host, m x k row-major).
// No Y_j may alias an S_i.
void block_combine(const double* const* S, int m, const double* C, double* const* Y, int k,
                   double beta, int64_t L, BlockWork& w, hipStream_t s);

// Symmetric eigendecomposition of a small dense matrix (cyclic Jacobi):
// A (n x n row-major, symmetric) -> w ascending, V[i n + j] = component i of
// eigenvector j.
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V);

// Smallest (which = 0) or largest (which = 1) eigenpairs of J(psi) at the
// parameters p, psi in internal order. evecs (nullable): k columns of 2n in
// internal order. Returns the number of converged eigenpairs.
int lobpcg_jacobian(nosh_nls_s* nls, const double* psi_int, const std::map<std::string, double>& p,
                    const nosh_eigen_options& o, double* evals, DBuf<double>* evecs,
                    nosh_eigen_stats* st, hipStream_t s);

}  // namespace nosh
