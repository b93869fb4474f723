// Finite-element mass and stiffness of the texel basis on the finest level.
//
// Output layout (the operator the multigrid hierarchy coarsens):
//   deep rows      : 10 doubles per deep texel, slot 3*(dj+1)+(di+1) for neighbour (di,dj),
//                    slot 9 unused (MassAndStiffness.h:288-294)
//   boundary rows  : A_BB (boundary x boundary) and A_BD (boundary x all texels, deep columns)
//                    as CSR, mass and stiffness kept separately (MassAndStiffness.h:490-549)
#pragma once
#include "gdtp_hierarchy.h"

namespace gdtp {

enum MetricMode { METRIC_EMBEDDING = 0, METRIC_UNIFORM = 1 };

struct FineOperator {
  std::vector<double> deepMass, deepStiffness;  // 10 * numDeep
  Csr bbMass, bbStiffness, bdMass, bdStiffness;
};

// Per chart, per chart triangle: the 2x2 metric of the parameterisation in UV coordinates,
// area-normalised so the surface has unit total area (Metric.inl:193-276, 415-446).
void ComputeParameterMetric(const TexturedMesh& mesh, const Atlas& atlas, int mode, std::vector<std::vector<M2>>& metric);

// Optional per-mesh-triangle metric override (e.g. anisotropic LIC metrics, Metric.inl:12-67):
// `surfaceMetric[t]` is the metric in the triangle's edge basis (the embedding Gram matrix).
void ComputeParameterMetricFromSurface(const TexturedMesh& mesh, const Atlas& atlas, const std::vector<M2>& surfaceMetric,
                                       std::vector<std::vector<M2>>& metric);

// Clip a convex polygon (chart-local coordinates) to the primal cell [i cw,(i+1) cw] x [j ch,(j+1) ch];
// false if nothing is left (ClipTriangleToPrimalCell, PolygonClipping.h:85-118).
bool ClipToCell(std::vector<V2>& poly, int i, int j, double cw, double ch);

// Assemble mass and stiffness over interior and boundary cells (MassAndStiffness.h:5-549).
int AssembleFineOperator(const TexturedMesh& mesh, const Hierarchy& h, const std::vector<std::vector<M2>>& metric,
                         FineOperator& op, std::string& err);

}  // namespace gdtp
