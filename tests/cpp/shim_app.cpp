// A TextureFiltering-shaped program written against include/gdtp_reference.hpp: the reference
// apps' call sites (InitializeHierarchy, UpdateLinearSystem, VCycle, solve,
// MultiplyBySystemMatrix_NoReciprocals; TextureFiltering.cpp:530-579, 581-718) with local
// stand-ins for the app's mesh and texel types.  Used by tests/test_cpp_shim.py.
//
//   shim_app host <mesh.bin> <width> <levels> <out.bin>   hierarchy only: texel atlas coordinates
//   shim_app gpu  <mesh.bin> <width> <levels> <out.bin>   + filtering: rhs = M f, x = f, 5 V-cycles,
//                                                          and the exact solution
// mesh.bin: int nv, int nt, double v[3 nv], int t[3 nt], double uv[6 nt] (per-corner UVs).
// f: texel i, channel c = ((i * 7 + c * 13) % 101) / 100.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gdtp_reference.hpp"

struct Point3D {
  double v[3] = {0, 0, 0};
  double& operator[](int i) { return v[i]; }
  const double& operator[](int i) const { return v[i]; }
};
struct Point2D {
  double v[2] = {0, 0};
  double& operator[](int i) { return v[i]; }
  const double& operator[](int i) const { return v[i]; }
};
struct TriangleIndex {
  int v[3] = {0, 0, 0};
  int& operator[](int i) { return v[i]; }
  const int& operator[](int i) const { return v[i]; }
};
struct TexturedMesh {  // the reference's SimpleMesh.h:45-49 shape
  std::vector<Point3D> vertices;
  std::vector<TriangleIndex> triangles;
  std::vector<Point2D> textureCoordinates;
};
struct TextureNodeInfo {  // Basis.h:4-30 fields used by the apps
  int ci = -1, cj = -1, chartId = -1;
};
struct CellIndex {};
struct AtlasChart {};
struct MultigridBlockInfo {};
struct SparseMatrixStandIn {};

static bool ReadMesh(const char* path, TexturedMesh& m) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  int nv = 0, nt = 0;
  bool ok = std::fread(&nv, 4, 1, f) == 1 && std::fread(&nt, 4, 1, f) == 1;
  std::vector<double> v(3 * (size_t)nv), uv(6 * (size_t)nt);
  std::vector<int> t(3 * (size_t)nt);
  ok = ok && std::fread(v.data(), 8, v.size(), f) == v.size() && std::fread(t.data(), 4, t.size(), f) == t.size() &&
       std::fread(uv.data(), 8, uv.size(), f) == uv.size();
  std::fclose(f);
  m.vertices.resize(nv), m.triangles.resize(nt), m.textureCoordinates.resize(3 * (size_t)nt);
  for (int i = 0; i < nv; i++)
    for (int c = 0; c < 3; c++) m.vertices[i][c] = v[3 * (size_t)i + c];
  for (int i = 0; i < nt; i++)
    for (int c = 0; c < 3; c++) m.triangles[i][c] = t[3 * (size_t)i + c];
  for (int i = 0; i < 3 * nt; i++)
    for (int c = 0; c < 2; c++) m.textureCoordinates[i][c] = uv[2 * (size_t)i + c];
  return ok;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: shim_app host|gpu mesh.bin width levels out.bin\n");
    return 2;
  }
  const bool gpu = std::strcmp(argv[1], "gpu") == 0;
  const int width = std::atoi(argv[3]), levels = std::atoi(argv[4]);
  TexturedMesh mesh;
  if (!ReadMesh(argv[2], mesh)) return 3;
  std::vector<TextureNodeInfo> textureNodes;
  std::vector<CellIndex> cellNodes;
  std::vector<AtlasChart> atlasCharts;
  gdtp_ref::HierarchicalSystem hierarchy;
  if (!gdtp_ref::InitializeHierarchy(mesh, width, width, levels, textureNodes, cellNodes, hierarchy, atlasCharts, MultigridBlockInfo(),
                                     true, false))
    return 4;
  FILE* out = std::fopen(argv[5], "wb");
  if (!out) return 5;
  const int N = (int)textureNodes.size();
  std::fwrite(&N, 4, 1, out);
  for (const TextureNodeInfo& t : textureNodes) {
    const int rec[3] = {t.ci, t.cj, t.chartId};
    std::fwrite(rec, 4, 3, out);
  }
  if (gpu) {
    // TextureFilter::InitializeSystem / UpdateSolution with Real = double, DataType = Point3D
    std::vector<gdtp_ref::MultigridLevelCoefficients<double>> coefficients;
    std::vector<gdtp_ref::MultigridLevelIndices<double>> indices(levels);
    std::vector<gdtp_ref::MultigridLevelVariables<Point3D>> variables(levels);
    std::vector<gdtp_ref::BoundarySolver> boundarySolver;
    gdtp_ref::CoarseSolver coarseSolver;
    gdtp_ref::DirectSolver directSolver;
    std::vector<double> deepMass, deepStiffness;
    SparseMatrixStandIn bbM, bbS, bdM, bdS, fine;
    if (!gdtp_ref::UpdateLinearSystem(1e-3, 1.0, hierarchy, coefficients, deepMass, deepStiffness, bbM, bbS, bdM, bdS, coarseSolver,
                                      boundarySolver, directSolver, fine, false, true, true))
      return 6;
    std::vector<Point3D> f(N), mf;
    for (int i = 0; i < N; i++)
      for (int c = 0; c < 3; c++) f[i][c] = ((i * 7 + c * 13) % 101) / 100.0;
    if (!gdtp_ref::MultiplyBySystemMatrix_NoReciprocals(coefficients, GDTP_OPERATOR_MASS, f, mf)) return 7;
    variables[0].x = f;
    variables[0].rhs = mf;
    for (int i = 0; i < 5; i++)
      if (!gdtp_ref::VCycle(variables, coefficients, indices, boundarySolver, coarseSolver, false, false)) return 8;
    std::vector<Point3D> exact;
    if (!gdtp_ref::solve(directSolver, exact, variables[0].rhs, coefficients)) return 9;
    for (const std::vector<Point3D>* v : {&mf, &variables[0].x, &exact})
      for (const Point3D& p : *v) std::fwrite(p.v, 8, 3, out);
  }
  std::fclose(out);
  std::printf("shim_app %s: %d texels\n", argv[1], N);
  return 0;
}
