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

// Cell-based stiffness right-hand side of the sharpening brush (MassAndStiffness.h:296-395,
// IterativeSolvers.inl:1170-1208): for signal columns f_k, every active cell's stiffness applied to
// f_k, kept per (texel, cell), so that per-cell weights w give t_k = sum_c w_c S_c f_k (w = 1:
// t_k = S f_k).  Deep texels have exactly their four interior cells; boundary texels a sparse row.
struct CellStiffnessRhs {
  int K = 0, numCells = 0;
  std::vector<double> deep;            // [K][4 numDeep]: deep texel d, cell slot s at 4d + s; slots
                                       // (i-1,j-1), (i,j-1), (i-1,j), (i,j) around texel (i,j)
  std::vector<int> deepCells;          // [2 numDeep]: global cells (i-1,j-1) and (i-1,j); +1 gives (i,j-1), (i,j)
  std::vector<int> bRowptr{0}, bCell;  // boundary rows (boundary index) -> global cells
  std::vector<double> bVal;            // [K][nnz]
};

// Assemble mass and stiffness over interior and boundary cells (MassAndStiffness.h:5-549).  With
// `cellRhs`, also the cell-based stiffness products of the K signal columns `signal` (host, [K][N]
// in global texel order).
int AssembleFineOperator(const TexturedMesh& mesh, const Hierarchy& h, const std::vector<std::vector<M2>>& metric,
                         FineOperator& op, std::string& err, const double* signal = nullptr, int K = 0,
                         CellStiffnessRhs* cellRhs = nullptr, bool deviceInterior = false);

// Interior cells on the GPU (gdtp_assembly_device.hip): the deep stencils (10 per deep texel, as
// above) and the 4x4 mass / stiffness matrices of the interior cells listed in `wanted` (global
// interior-cell indices: charts in order, each chart's interior cells in raster order), 16 per cell.
void AssembleInteriorCellsOnDevice(const Hierarchy& h, const std::vector<std::vector<M2>>& metric, const std::vector<int>& wanted,
                                   std::vector<double>& deepMass, std::vector<double>& deepStiffness, std::vector<double>& wantedM,
                                   std::vector<double>& wantedS);

}  // namespace gdtp
