// Multigrid operator hierarchy for A = screen*M + stiff*S (Galerkin coarsening) and the exact
// boundary / coarse factorisations; the host half of UpdateLinearSystem
// (include/Src/UpdateCoefficients.inl:285-513).
#pragma once
#include "gdtp_assembly.h"
#include "gdtp_ldlt.h"

namespace gdtp {

struct LevelOperator {
  std::vector<double> deepRaw;    // 10 per deep texel, A entries (slot 9 = 0)
  std::vector<double> deepRecip;  // reciprocal form: c_k = a_k/a_4 (k != 4), c_4 = 1/a_4, c_9 = a_4
  Csr bd, bb;                     // boundary x all texels (deep columns), boundary x boundary
};

struct OperatorHierarchy {
  std::vector<LevelOperator> levels;
  Csr coarseMatrix;                     // full matrix of the coarsest level
  std::vector<Ldlt> boundaryFactor;     // levels 0 .. L-2
  Ldlt coarseFactor;
};

// Galerkin R A P of a 9-point deep stencil field along the coarse deep lines
// (DeepCoefficientRestriction, UpdateCoefficients.inl:1-87).
void RestrictDeepCoefficients(const std::vector<double>& fine, std::vector<double>& coarse, const std::vector<DeepLine>& lines);

// Full sparse matrix of a level from deep rows + boundary rows (FullMatrixConstruction, :137-219).
Csr FullLevelMatrix(const Level& L, const std::vector<double>& deepRaw, const Csr& bb, const Csr& bd);

int BuildOperatorHierarchy(const Hierarchy& h, const FineOperator& op, double screen, double stiff, bool factor,
                           OperatorHierarchy& oh, std::string& err);

}  // namespace gdtp
