// Sparse LDL^T factorisation (host) of symmetric positive-definite matrices, with a
// nested-dissection ordering so that the triangular solves have shallow dependency levels.
//
// Replaces the reference's exact boundary and coarse solvers, Eigen SimplicialLDLT
// (include/Src/Solver.h:145-259).  Factorisation: up-looking LDL^T (T. A. Davis, "Algorithm 849:
// a concise sparse Cholesky factorization package", ACM TOMS 31(4), 2005).  Ordering: recursive
// breadth-first-level bisection (A. George, "Nested dissection of a regular finite element mesh",
// SIAM J. Numer. Anal. 10, 1973).
#pragma once
#include <string>

#include "gdtp_sparse.h"

namespace gdtp {

struct Ldlt {
  int n = 0;
  std::vector<int> perm;      // new index -> old index
  std::vector<int> iperm;     // old index -> new index
  Csr Lrows;                  // strictly lower triangular L by rows (row k: entries l_ki, i < k), permuted space
  Csr Lcols;                  // the same entries by columns (row of L^T)
  std::vector<double> D;      // diagonal
  // level schedules: rows grouped by dependency level, forward (L) and backward (L^T)
  std::vector<int> fwdLevelPtr, fwdRows, bwdLevelPtr, bwdRows;
};

void NestedDissectionOrder(const Csr& A, std::vector<int>& perm);
int FactorLdlt(const Csr& A, Ldlt& F, std::string& err);
// Solve A x = b for ncols column-major vectors (host reference path).
void SolveLdlt(const Ldlt& F, const double* b, double* x, int ncols, int ld);

}  // namespace gdtp
