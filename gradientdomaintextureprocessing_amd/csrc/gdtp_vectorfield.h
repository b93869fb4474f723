// Samples for integrating a normalised gradient field against the texel basis (the divergence
// right-hand side of the heat-method geodesics; VectorFieldIntergration.inl:71-582).
//
// For a potential u on the texels, the output is, per texel t,
//     rhs_t = sum over samples s of  < X_s , v_s(t) >,   X_s = g^-1 N / sqrt(N^T g^-1 N),  N = -grad u(s)
// i.e. the one-point-per-piece quadrature of  integral < -grad u / |grad u|_g , grad phi_t >_g dA.
#pragma once
#include "gdtp_boundary.h"

namespace gdtp {

struct VectorFieldSamples {
  // Cell samples: one per (mesh triangle ∩ cell) piece away from seams; the field there is the
  // bilinear interpolant of the 4 cell corners (global texel indices, corner order
  // (0,0),(1,0),(1,1),(0,1)).  Data per sample (FP32): position in the unit cell (x, y), the
  // inverse metric (t00, t01, t11), and v_k = sqrt(det g) * integral over the piece of grad phi_k
  // (UV units) for k = 0..3 (x, y each): 13 floats.
  std::vector<int> cellCorners;
  std::vector<float> cellData;
  double invCellW = 0, invCellH = 0;
  // General samples: one per quadratic triangle of a seam piece, evaluated at its barycentre.
  // Entries (global texel, G = d(grad u)/d(u_texel), V = the texel's integrated basis gradient).
  std::vector<int> genPtr{0}, genTexel;
  std::vector<float> genTensor, genG, genV;  // 3 per sample; 2 per entry; 2 per entry
  size_t numCellSamples() const { return cellCorners.size() / 4; }
  size_t numGeneralSamples() const { return genPtr.size() - 1; }
};

int BuildVectorFieldSamples(const Hierarchy& h, const std::vector<std::vector<M2>>& metric, VectorFieldSamples& out,
                            std::string& err);

}  // namespace gdtp
