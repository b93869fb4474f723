// Boundary-cell and seam assembly of the finest-level operator.
#pragma once
#include "gdtp_assembly.h"

namespace gdtp {

// Boundary cells of chart `c`: bilinear texel basis integrated exactly over (cell ∩ chart),
// accumulated per boundary cell in bcellM / bcellS (4x4, corner order (0,0),(1,0),(1,1),(0,1))
// and scattered into boundary-boundary triplets.
int AssembleBoundaryCells(const TexturedMesh& mesh, const Hierarchy& h, int c, const std::vector<M2>& metric,
                          const std::vector<std::array<double, 16>>& bcellM, const std::vector<std::array<double, 16>>& bcellS,
                          std::vector<Triplet>& bbM, std::vector<Triplet>& bbS, std::string& err);

// Cross-chart coupling along seams.
int AssembleSeams(const TexturedMesh& mesh, const Hierarchy& h, const std::vector<std::vector<M2>>& metric, FineOperator& op,
                  std::string& err);

}  // namespace gdtp
