// Texel hierarchy of a texture atlas: charts, per-level texel grids, texel classification,
// raster-order indexing, Gauss-Seidel band tasks and inter-level transfer lines.
//
// Semantics follow the reference hierarchy builder:
//   atlas charts                 include/Src/ChartDecomposition.h:30-73, AtlasMesh.inl:3-184,
//                                AtlasCharts.inl:3-107
//   per-level grid charts        include/Src/HierarchyConstruction.inl:496-532
//   texel/cell classification,
//   raster lines, band tasks     include/Src/HierarchyConstruction.inl:39-493
//   boundary/deep indexing       include/Src/HierarchyConstruction.inl:3-35
//   restriction / prolongation   include/Src/ProlongationAndRestriction.inl:99-452
//   Galerkin boundary indices    include/Src/HierarchyConstruction.inl:568-706
#pragma once
#include <array>
#include <unordered_map>

#include "gdtp_common.h"
#include "gdtp_sparse.h"

namespace gdtp {

// ------------------------------------------------------------------ atlas
struct AtlasChart {
  V2 minCorner, maxCorner, gridOrigin;
  int originCoords[2];
  std::vector<int> triangles;             // chart-local vertex ids, 3 per triangle
  std::vector<V2> vertices;               // chart-local UV positions
  std::vector<int> boundaryHalfEdges;     // chart-local half-edge ids (3*t + k)
  std::vector<int> atlasEdgeIndices;      // 3 per chart triangle
  std::vector<int> meshVertexIndices;     // per chart vertex
  std::vector<int> meshTriangleIndices;   // per chart triangle
};

struct Atlas {
  int numCharts = 0;
  std::vector<int> triangleChart;         // per mesh triangle
  std::vector<int> oppositeHalfEdge;      // per mesh half-edge, -1 on open boundary
  std::vector<int> halfEdgeToEdgeIndex;   // per mesh half-edge (atlas edge ids)
  std::unordered_map<int, int> boundaryVertexIndex;  // mesh vertex -> boundary vertex id
  int numBoundaryVertices = 0;
  bool isClosed = true;
  std::vector<AtlasChart> charts;
};

int BuildAtlas(const TexturedMesh& mesh, int W, int H, Atlas& atlas, std::string& err);

// ------------------------------------------------------------------ hierarchy
struct Segment {  // one run of deep texels in a raster row (global / deep indices)
  int currentStart, currentEnd, previousStart, nextStart, deepStart;
};

struct GridChart {
  V2 corner;
  int cornerCoords[2], centerOffset[2];
  double cellW, cellH;
  int width, height;
  int texelOffset, interiorTexelOffset, deepOffset, boundaryOffset, cellOffset, gridIndexOffset;
  int numInteriorCells = 0, numBoundaryCells = 0;
  // node type: -1 inactive, 0 exterior, 1 interior boundary, 2 deep.  cell type: -1, 0 boundary, 1 interior
  std::vector<int8_t> nodeType, cellType;
  std::vector<int> globalTexel, globalInterior, globalDeep, globalBoundary;  // width*height, -1 if none
  std::vector<int> localCell, localInteriorCell, localBoundaryCell;          // (width-1)*(height-1)
  std::vector<std::array<int, 4>> cellIndices;                // active cells: global texel of corners (ccw from (i,j))
  std::vector<std::array<int, 4>> interiorCellCorners;        // interior cells: global interior-texel index of corners
  std::vector<std::array<int, 4>> interiorCellGlobalCorners;  // interior cells: global texel index of corners
  std::vector<int> interiorCellToLocal, boundaryCellToLocal;
  int node(int i, int j) const { return nodeType[(size_t)j * width + i]; }
  int cell(int i, int j) const { return cellType[(size_t)j * (width - 1) + i]; }
  int texel(int i, int j) const { return globalTexel[(size_t)j * width + i]; }
  int deep(int i, int j) const { return globalDeep[(size_t)j * width + i]; }
};

struct NodeInfo { int chart, ci, cj, type; };

struct Line5 { int start, end, prev, next, coeff; };
struct DeepLine { int coarseStart, coarseEnd, finePrev, fineCurrent, fineNext; };
struct ProlongationLine { int start, length, center, prev, next, aligned; };

// Flattened band tasks: task t owns blocks [taskPtr[t], taskPtr[t+1]); block b owns deep rows
// [lineBegin[b], deepEnd[b]) then padded rows [deepEnd[b], padEnd[b]); row r owns segments
// [linePtr[r], linePtr[r+1]).
// Wave groups of the GPU band kernel: blocks of one band task in flight at once.
constexpr int kBandGroups = 3;

struct BandTasks {
  std::vector<int> taskPtr{0}, blockBegin, blockDeepEnd, blockPadEnd, linePtr{0};
  std::vector<Segment> segments;
  std::vector<int> lineRow;    // chart grid row of every line (deep and padded)
  std::vector<int> blockStart; // GPU pipeline: first task step of every block (BandSchedule)
  std::vector<int> taskSteps;  // GPU pipeline: task steps of every task
  // GPU row elements: line l's elements (its segments' texels in order) are lineElem[l] ..
  // lineElem[l + 1] - 1, element e stored as 4 ints (global index, index above, index below, deep
  // index d; -d - 1 where the element starts a segment)
  std::vector<int> lineElem, elems;
};

// The GPU band kernel's pipeline schedule and row element table of T (fills blockStart, taskSteps,
// lineElem and elems).
void BandSchedule(BandTasks& T);

struct Level {
  std::vector<GridChart> charts;
  int numTexels = 0, numInterior = 0, numDeep = 0, numBoundary = 0, numCells = 0, numBoundaryCells = 0, numInteriorCells = 0;
  std::vector<int> boundaryAndDeepIndex, boundaryGlobalIndex, deepGlobalIndex;
  std::vector<NodeInfo> nodeInfo;
  std::vector<Line5> rasterLines;
  BandTasks tasks;
  std::vector<Line5> restrictionLines;      // coarse levels: fine rows feeding each coarse deep run
  std::vector<DeepLine> deepLines;          // coarse levels: Galerkin deep coefficient rows
  std::vector<ProlongationLine> prolongationLines;  // fine levels: all active texels by row runs
  Csr boundaryRestriction;                  // fine levels: (coarse boundary) x (fine texels)
};

struct BlockInfo { int blockWidth = 128, blockHeight = 16, paddingWidth = 2, paddingHeight = 0; };

struct Hierarchy {
  int width = 0, height = 0;
  Atlas atlas;
  std::vector<Level> levels;
};

int BuildHierarchy(const TexturedMesh& mesh, int W, int H, int numLevels, const BlockInfo& bi, Hierarchy& h, std::string& err);

}  // namespace gdtp
