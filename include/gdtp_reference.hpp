// gdtp_reference.hpp — the reference apps' C++ call surface over the C ABI of libgdtp.
//
// TextureFiltering, Geodesics and LineIntegralConvolution call header templates directly
// (InitializeHierarchy, UpdateLinearSystem, VCycle, solve, MultiplyBySystemMatrix_NoReciprocals).
// These templates keep those names, argument lists and meanings, and the reference's error
// convention (return 1 on success; on failure print "ERROR: ..." and return 0), so the apps' call
// sites compile unchanged.  What changes is the data behind the arguments: the per-level
// coefficient, index and solver objects become handles of libgdtp (the hierarchy, the operator
// hierarchy with its factorisations, and device buffers), declared here under the reference's
// type names in namespace gdtp_ref.  Texel vectors stay the apps' std::vector of scalars or of
// three-channel values (anything with operator[] for channels 0..2), in the reference's global
// texel order; each call stages them through the device once.
//
//   reference (file:line)                               here
//   InitializeHierarchy (HierarchyConstruction.inl:799)  gdtp_mesh_create + gdtp_hierarchy_create
//   UpdateLinearSystem (UpdateCoefficients.inl:463-469)  gdtp_system_create / gdtp_system_update
//   VCycle (IterativeSolvers.inl:1211-1216)              gdtp_vcycle
//   solve(directSolver, ...) (Solver.h:191-259)          gdtp_direct_solve
//   MultiplyBySystemMatrix_NoReciprocals (:881)          gdtp_apply (A, M or S)
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <type_traits>
#include <vector>

#include "gdtp.h"

namespace gdtp_ref {

// Channels of a texel value: 1 for arithmetic scalars, else 3 (Point3D<Real>, Vec4 lanes 0..2).
template <class T>
constexpr int ChannelsOf() {
  return std::is_arithmetic<T>::value ? 1 : 3;
}
template <class T>
double Channel(const T& v, int c) {
  if constexpr (std::is_arithmetic<T>::value) return (void)c, static_cast<double>(v);
  else return static_cast<double>(v[c]);
}
template <class T>
void SetChannel(T& v, int c, double x) {
  if constexpr (std::is_arithmetic<T>::value) (void)c, v = static_cast<T>(x);
  else v[c] = x;
}

inline int Report(int rc) {
  if (rc != GDTP_OK) std::printf("ERROR: %s\n", gdtp_last_error());
  return rc == GDTP_OK ? 1 : 0;
}

// The texel hierarchy of one atlas (replaces the reference's HierarchicalSystem).
struct HierarchicalSystem {
  gdtp_mesh* mesh = nullptr;
  gdtp_hierarchy* hierarchy = nullptr;
  int levels = 0, texels = 0;
  HierarchicalSystem() = default;
  HierarchicalSystem(const HierarchicalSystem&) = delete;
  HierarchicalSystem& operator=(const HierarchicalSystem&) = delete;
  ~HierarchicalSystem() {
    gdtp_hierarchy_destroy(hierarchy);
    gdtp_mesh_destroy(mesh);
  }
};

// Per-level texel vectors, the reference's MultigridLevelVariables (Hierarchy.h:287-296); only
// level 0's x and rhs are read and written here, the other levels live on the device.
template <class DataType>
struct MultigridLevelVariables {
  std::vector<DataType> x, rhs, residual, boundary_rhs, boundary_value, variable_boundary_value;
};

// The operator hierarchy A = screen M + stiffness S with its factorisations and device buffers
// (replaces the reference's MultigridLevelCoefficients, one per level: the handle sits in the
// first entry of the vector).
template <class Real>
struct MultigridLevelCoefficients {
  gdtp_system* system = nullptr;
  int columns = 0, texels = 0;
  double *dx = nullptr, *db = nullptr;  // device staging, [columns][texels]
  MultigridLevelCoefficients() = default;
  MultigridLevelCoefficients(const MultigridLevelCoefficients&) = delete;
  MultigridLevelCoefficients& operator=(const MultigridLevelCoefficients&) = delete;
  MultigridLevelCoefficients(MultigridLevelCoefficients&& o) noexcept { *this = static_cast<MultigridLevelCoefficients&&>(o); }
  MultigridLevelCoefficients& operator=(MultigridLevelCoefficients&& o) noexcept {
    std::swap(system, o.system), std::swap(columns, o.columns), std::swap(texels, o.texels), std::swap(dx, o.dx), std::swap(db, o.db);
    return *this;
  }
  ~MultigridLevelCoefficients() {
    if (dx) (void)hipFree(dx);
    if (db) (void)hipFree(db);
    gdtp_system_destroy(system);
  }
};

// The hierarchy's index structures of one level (MultigridLevelIndices); held by the hierarchy.
template <class Real>
struct MultigridLevelIndices {
  const HierarchicalSystem* hierarchy = nullptr;
  int level = 0;
};

// The exact boundary / coarse solvers and the direct solver live inside the operator hierarchy;
// these carry only the handle they solve with.
struct BoundarySolver {};
struct CoarseSolver {};
struct DirectSolver {
  gdtp_system* system = nullptr;
};

// InitializeHierarchy (HierarchyConstruction.inl:799): mesh -> texel hierarchy.  Mesh: the
// reference's TexturedMesh shape (vertices[v][0..2], triangles[t][0..2], textureCoordinates[3 t + k]
// [0..1]).  textureNodes receive, per texel in global order, the atlas column / row (ci, cj) and
// the chart (chartId); cellNodes, atlasCharts and blockInfo are the reference's outputs that the
// library keeps internally and are left untouched.
template <class Mesh, class TextureNode, class CellNode, class AtlasChart, class BlockInfo>
int InitializeHierarchy(Mesh& mesh, const int width, const int height, const int levels, std::vector<TextureNode>& textureNodes,
                        std::vector<CellNode>& cellNodes, HierarchicalSystem& hierarchy, std::vector<AtlasChart>& atlasCharts,
                        const BlockInfo& blockInfo, bool verbose = false, bool detailVerbose = false, bool computeProlongation = false) {
  (void)cellNodes, (void)atlasCharts, (void)blockInfo, (void)detailVerbose, (void)computeProlongation;
  const int nv = (int)mesh.vertices.size(), nt = (int)mesh.triangles.size();
  std::vector<double> v(3 * (size_t)nv), uv(6 * (size_t)nt);
  std::vector<int> t(3 * (size_t)nt);
  for (int i = 0; i < nv; i++)
    for (int c = 0; c < 3; c++) v[3 * (size_t)i + c] = mesh.vertices[i][c];
  for (int i = 0; i < nt; i++)
    for (int c = 0; c < 3; c++) t[3 * (size_t)i + c] = mesh.triangles[i][c];
  for (int i = 0; i < 3 * nt; i++)
    for (int c = 0; c < 2; c++) uv[2 * (size_t)i + c] = mesh.textureCoordinates[i][c];
  if (!Report(gdtp_mesh_create(v.data(), nv, t.data(), nt, uv.data(), &hierarchy.mesh))) return 0;
  if (!Report(gdtp_hierarchy_create(hierarchy.mesh, width, height, levels, GDTP_METRIC_EMBEDDING, &hierarchy.hierarchy))) return 0;
  int info[8];
  if (!Report(gdtp_hierarchy_level_info(hierarchy.hierarchy, 0, info))) return 0;
  hierarchy.levels = levels, hierarchy.texels = info[0];
  size_t nNode = 0, nChart = 0;
  if (!Report(gdtp_hierarchy_array_i32(hierarchy.hierarchy, 0, "nodeInfo", nullptr, &nNode)) ||
      !Report(gdtp_hierarchy_array_i32(hierarchy.hierarchy, 0, "charts", nullptr, &nChart)))
    return 0;
  std::vector<int> node(nNode), chart(nChart);
  gdtp_hierarchy_array_i32(hierarchy.hierarchy, 0, "nodeInfo", node.data(), &nNode);
  gdtp_hierarchy_array_i32(hierarchy.hierarchy, 0, "charts", chart.data(), &nChart);
  textureNodes.assign(info[0], TextureNode());
  for (int i = 0; i < info[0]; i++) {  // nodeInfo: chart, i, j, type; charts: 9 ints, corner at 2..3
    const int c = node[4 * (size_t)i];
    textureNodes[i].chartId = c;
    textureNodes[i].ci = chart[9 * (size_t)c + 2] + node[4 * (size_t)i + 1];
    textureNodes[i].cj = chart[9 * (size_t)c + 3] + node[4 * (size_t)i + 2];
  }
  if (verbose) std::printf("Hierarchy: %d levels, %d texels\n", levels, info[0]);
  return 1;
}

// UpdateLinearSystem (UpdateCoefficients.inl:463-469): A = screenWeight M + stiffnessWeight S on
// every level, the exact boundary / coarse factorisations and, with updateFineSolver, the direct
// solver.  M and S are the library's assembly of the same mesh, so the app's mass / stiffness
// arguments and fine system matrix are not read.  The first call creates the operator hierarchy
// with device buffers for `columns` texel channels (3 by default, as for RGB).
template <class Real, class DeepCoefficients, class SparseMatrixType, class CoarseSolverType, class BoundarySolverType,
          class DirectSolverType>
int UpdateLinearSystem(double stiffnessWeight, double screenWeight, const HierarchicalSystem& hierarchy,
                       std::vector<MultigridLevelCoefficients<Real>>& multigridCoefficients, const DeepCoefficients& deepMassCoefficients,
                       const DeepCoefficients& deepStiffnessCoefficients, const SparseMatrixType& boundaryBoundaryMassMatrix,
                       const SparseMatrixType& boundaryBoundaryStiffnessMatrix, const SparseMatrixType& boundaryDeepMassMatrix,
                       const SparseMatrixType& boundaryDeepStiffnessMatrix, CoarseSolverType& coarseSolver,
                       std::vector<BoundarySolverType>& boundarySolver, DirectSolverType& fineSolver,
                       const SparseMatrixType& fineSystemMatrix, bool detailVerbose = false, bool initSolvers = true,
                       bool updateFineSolver = false, int columns = 3) {
  (void)deepMassCoefficients, (void)deepStiffnessCoefficients, (void)boundaryBoundaryMassMatrix, (void)boundaryBoundaryStiffnessMatrix;
  (void)boundaryDeepMassMatrix, (void)boundaryDeepStiffnessMatrix, (void)coarseSolver, (void)fineSystemMatrix, (void)detailVerbose;
  (void)initSolvers;
  if (multigridCoefficients.size() < (size_t)hierarchy.levels) multigridCoefficients.resize(hierarchy.levels);
  boundarySolver.resize(hierarchy.levels > 0 ? hierarchy.levels - 1 : 0);
  MultigridLevelCoefficients<Real>& c = multigridCoefficients[0];
  if (!c.system) {
    if (!Report(gdtp_system_create(hierarchy.hierarchy, screenWeight, stiffnessWeight, columns, &c.system))) return 0;
    c.columns = columns, c.texels = hierarchy.texels;
    const size_t bytes = sizeof(double) * (size_t)columns * c.texels;
    if (hipMalloc((void**)&c.dx, bytes) != hipSuccess || hipMalloc((void**)&c.db, bytes) != hipSuccess) {
      std::printf("ERROR: device staging allocation failed\n");
      return 0;
    }
  } else if (!Report(gdtp_system_update(c.system, screenWeight, stiffnessWeight))) {
    return 0;
  }
  if (updateFineSolver) fineSolver.system = c.system;
  return 1;
}

// Level-0 vectors of the app (columns = channels) to / from the device staging buffers.
template <class Real, class DataType>
int Stage(MultigridLevelCoefficients<Real>& c, const std::vector<DataType>& v, double* dev) {
  constexpr int K = ChannelsOf<DataType>();
  if ((int)v.size() != c.texels || K > c.columns) {
    std::printf("ERROR: %zu texel values of %d channels for a system of %d texels and %d columns\n", v.size(), K, c.texels, c.columns);
    return 0;
  }
  std::vector<double> h((size_t)K * c.texels);
  for (int i = 0; i < c.texels; i++)
    for (int k = 0; k < K; k++) h[(size_t)k * c.texels + i] = Channel(v[i], k);
  return hipMemcpy(dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
}
template <class Real, class DataType>
int Unstage(MultigridLevelCoefficients<Real>& c, const double* dev, std::vector<DataType>& v) {
  constexpr int K = ChannelsOf<DataType>();
  std::vector<double> h((size_t)K * c.texels);
  if (hipMemcpy(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 0;
  v.resize(c.texels);
  for (int i = 0; i < c.texels; i++)
    for (int k = 0; k < K; k++) SetChannel(v[i], k, h[(size_t)k * c.texels + i]);
  return 1;
}

// VCycle (IterativeSolvers.inl:1211-1216): one V-cycle of the reference schedule on level 0,
// variables[0].x in / out, variables[0].rhs in, one column per channel.
template <class Real, class DataType, class BoundarySolverType, class CoarseSolverType>
int VCycle(std::vector<MultigridLevelVariables<DataType>>& variables, std::vector<MultigridLevelCoefficients<Real>>& coefficients,
           const std::vector<MultigridLevelIndices<Real>>& indices, std::vector<BoundarySolverType>& boundarySolvers,
           CoarseSolverType& coarseSolver, bool verbose = false, bool detailVerbose = false) {
  (void)indices, (void)boundarySolvers, (void)coarseSolver, (void)verbose, (void)detailVerbose;
  MultigridLevelCoefficients<Real>& c = coefficients.at(0);
  MultigridLevelVariables<DataType>& v = variables.at(0);
  if (!Stage(c, v.x, c.dx) || !Stage(c, v.rhs, c.db)) return 0;
  if (!Report(gdtp_vcycle(c.system, c.dx, c.db, ChannelsOf<DataType>(), nullptr))) return 0;
  return Unstage(c, c.dx, v.x);
}

// solve(directSolver, x, rhs) (Solver.h:191-259) for the direct solver of the apps'
// --useDirectSolver path: exact FP64 solve of the level-0 system.
template <class Real, class DataType>
int solve(DirectSolver& solver, std::vector<DataType>& x, const std::vector<DataType>& rhs,
          std::vector<MultigridLevelCoefficients<Real>>& coefficients) {
  MultigridLevelCoefficients<Real>& c = coefficients.at(0);
  if (solver.system != c.system) {
    std::printf("ERROR: direct solver not set up by UpdateLinearSystem(..., updateFineSolver = true)\n");
    return 0;
  }
  if (!Stage(c, rhs, c.db)) return 0;
  if (!Report(gdtp_direct_solve(c.system, c.dx, c.db, ChannelsOf<DataType>(), nullptr, nullptr, nullptr))) return 0;
  return Unstage(c, c.dx, x);
}

// MultiplyBySystemMatrix_NoReciprocals (IterativeSolvers.inl:881) with the level-0 operator of
// the system: which = GDTP_OPERATOR_SYSTEM (A), GDTP_OPERATOR_MASS (M) or GDTP_OPERATOR_STIFFNESS (S)
// (the apps form rhs = M f + w S f this way, TextureFiltering.cpp:692-701).
template <class Real, class DataType>
int MultiplyBySystemMatrix_NoReciprocals(std::vector<MultigridLevelCoefficients<Real>>& coefficients, int which,
                                         const std::vector<DataType>& in, std::vector<DataType>& out) {
  MultigridLevelCoefficients<Real>& c = coefficients.at(0);
  if (!Stage(c, in, c.dx)) return 0;
  if (!Report(gdtp_apply(c.system, which, c.dx, c.db, ChannelsOf<DataType>(), nullptr))) return 0;
  return Unstage(c, c.db, out);
}

}  // namespace gdtp_ref
