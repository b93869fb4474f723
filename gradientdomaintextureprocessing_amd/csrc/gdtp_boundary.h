// Boundary-cell assembly of the finest-level operator, with the chart-to-chart coupling along
// seams.
//
// The reference represents the texture on boundary cells (cells cut by the chart boundary) with
// quadratic elements whose nodal values are interpolated from the bilinear texel basis
// (BoundaryQuadraticElements.inl:652-773): a node takes the mean, over every chart it belongs
// to, of that chart's bilinear interpolation at the node.  Nodes on seams (shared mesh vertices,
// crossings of a seam edge with either chart's grid lines, midpoints of the seam segments
// between them) belong to both charts; every other node to one.  A quadratic interpolant of a
// bilinear function is the function itself, so a boundary piece with no seam node carries
// exactly the bilinear basis and is integrated as such (AssembleFineOperator); a piece that
// touches a seam is split into quadratic triangles whose seam nodes carry the averaged values.
#pragma once
#include "gdtp_assembly.h"

namespace gdtp {

// Seam topology of the atlas at level 0: seam half-edges, the union of both charts' grid-line
// crossings along each seam edge, and the charts sharing each seam vertex.
struct SeamContext {
  const Hierarchy* h = nullptr;
  std::vector<int> meshTriangleLocal;              // mesh triangle -> index within its chart
  std::vector<char> seamHalfEdge;                  // per mesh half-edge: chart boundary with an opposite
  std::vector<std::vector<double>> crossings;      // per seam mesh half-edge: sorted crossing parameters (0,1)
  std::vector<int> vertexChartCount;               // per mesh vertex: appearances as a chart-boundary vertex
  std::vector<std::vector<std::pair<int, int>>> vertexCharts;  // per mesh vertex: (chart, chart-local vertex)
};

int BuildSeamContext(const Hierarchy& h, SeamContext& ctx, std::string& err);

// (global boundary index, weight) pairs: the value of a node as a combination of texel values.
using NodeWeights = std::vector<std::pair<int, double>>;

// One quadratic triangle of a seam piece (chart-local coordinates); nodes: corners 0,1,2, then the
// midpoints of (0,1), (1,2), (2,0).
struct SeamElement {
  V2 p[3];
  NodeWeights w[6];
};

// The quadratic elements of a piece that touches a seam (1), none if it does not (0), -1 on error.
int SeamPieceElements(const SeamContext& ctx, int c, int t, int i, int j, const std::vector<V2>& piece, std::vector<SeamElement>& els,
                      std::string& err);

// Piece = (chart triangle t of chart c) ∩ (boundary cell (i,j)), as a convex polygon in chart-local
// coordinates.  If the piece touches a seam, integrate its quadratic elements and append the
// boundary-boundary triplets (global boundary indices) to bbM / bbS and return 1; return 0 if it
// does not touch a seam (the caller integrates the bilinear basis); -1 on error (err set).
int AssembleSeamPiece(const SeamContext& ctx, int c, int t, int i, int j, const std::vector<V2>& piece, const M2& metric,
                      std::vector<Triplet>& bbM, std::vector<Triplet>& bbS, std::string& err);

// Boundary cells of chart `c`: scatter the accumulated 4x4 bilinear cell matrices (corner order
// (0,0),(1,0),(1,1),(0,1)) into boundary-boundary triplets.
int AssembleBoundaryCells(const Hierarchy& h, int c, const std::vector<std::array<double, 16>>& bcellM,
                          const std::vector<std::array<double, 16>>& bcellS, std::vector<Triplet>& bbM, std::vector<Triplet>& bbS,
                          std::string& err);

}  // namespace gdtp
