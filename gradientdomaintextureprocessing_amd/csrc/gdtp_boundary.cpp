// Boundary-cell assembly.
//
// Inside one chart, the reference's quadratic boundary elements (BoundaryQuadraticElements.inl,
// MassAndStiffness.h:146-272) interpolate the bilinear texel basis at P2 nodes; P2 reproduces
// bilinear functions exactly, so away from seams the restricted boundary operator equals the
// exact integral of the bilinear basis over (cell ∩ chart), which is what this computes.
#include "gdtp_boundary.h"

namespace gdtp {

int AssembleBoundaryCells(const TexturedMesh& mesh, const Hierarchy& h, int c, const std::vector<M2>& metric,
                          const std::vector<std::array<double, 16>>& bcellM, const std::vector<std::array<double, 16>>& bcellS,
                          std::vector<Triplet>& bbM, std::vector<Triplet>& bbS, std::string& err) {
  (void)mesh, (void)metric;
  const Level& L0 = h.levels[0];
  const GridChart& g = L0.charts[c];
  for (int bc = 0; bc < g.numBoundaryCells; bc++) {
    const auto& gl = g.cellIndices[g.boundaryCellToLocal[bc]];
    int b[4];
    for (int k = 0; k < 4; k++) {
      b[k] = L0.boundaryAndDeepIndex[gl[k]] - 1;
      if (b[k] < 0) { err = "deep texel on a boundary cell"; return 0; }
    }
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < 4; l++) {
        bbM.push_back({b[k], b[l], bcellM[bc][4 * k + l]});
        bbS.push_back({b[k], b[l], bcellS[bc][4 * k + l]});
      }
  }
  return 1;
}

int AssembleSeams(const TexturedMesh& mesh, const Hierarchy& h, const std::vector<std::vector<M2>>& metric, FineOperator& op,
                  std::string& err) {
  (void)mesh, (void)h, (void)metric, (void)op, (void)err;
  return 1;
}

}  // namespace gdtp
