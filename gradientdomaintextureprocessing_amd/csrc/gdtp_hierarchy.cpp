// Atlas and texel-hierarchy construction (host side, once per mesh and resolution).
// Each function names the reference routine whose semantics it follows.
#include "gdtp_hierarchy.h"

#include <algorithm>
#include <cstdint>

#include <cstring>
#include <map>
#include <queue>

namespace gdtp {

static inline uint64_t EdgeKey(uint32_t a, uint32_t b) { return (uint64_t(a) << 32) | uint64_t(b); }

// ---------------------------------------------------------------- atlas
// Chart ids by breadth-first search over triangles whose shared edge has matching corner UVs
// (ChartDecomposition.h:30-73); atlas vertices are unique (uv, mesh vertex) pairs
// (AtlasMesh.inl:3-34); atlas edge ids (AtlasMesh.inl:68-96); seam / open-boundary half-edges
// and boundary vertex ids (AtlasMesh.inl:102-184); per-chart arrays and grid origins
// (AtlasCharts.inl:3-75).
int BuildAtlas(const TexturedMesh& mesh, int W, int H, Atlas& A, std::string& err) {
  const int nT = mesh.numTriangles();
  const auto& T = mesh.triangles;
  const auto& UV = mesh.uv;
  std::unordered_map<uint64_t, int> meshHalfEdge;
  meshHalfEdge.reserve(3 * (size_t)nT);
  for (int t = 0; t < nT; t++)
    for (int k = 0; k < 3; k++) {
      uint64_t key = EdgeKey(T[3 * t + k], T[3 * t + (k + 1) % 3]);
      if (!meshHalfEdge.emplace(key, 3 * t + k).second) { err = "non-manifold mesh"; return 0; }
    }
  auto sameUV = [&](int a, int b) { return UV[a].x == UV[b].x && UV[a].y == UV[b].y; };
  // chart decomposition
  std::vector<std::vector<int>> nbr(nT);
  for (int t = 0; t < nT; t++)
    for (int k = 0; k < 3; k++) {
      auto it = meshHalfEdge.find(EdgeKey(T[3 * t + (k + 1) % 3], T[3 * t + k]));
      if (it == meshHalfEdge.end()) continue;
      int o = it->second / 3, ok = it->second % 3;
      if (sameUV(3 * t + (k + 1) % 3, 3 * o + ok) && sameUV(3 * t + k, 3 * o + (ok + 1) % 3)) nbr[t].push_back(o);
    }
  A.triangleChart.assign(nT, -1);
  int comp = -1;
  for (int s = 0; s < nT; s++) {
    if (A.triangleChart[s] != -1) continue;
    comp++;
    std::queue<int> q;
    A.triangleChart[s] = comp;
    q.push(s);
    while (!q.empty()) {
      int c = q.front();
      q.pop();
      for (int n : nbr[c]) {
        if (A.triangleChart[n] == -1) A.triangleChart[n] = comp, q.push(n);
        else if (A.triangleChart[n] != comp) { err = "inconsistent chart components"; return 0; }
      }
    }
  }
  A.numCharts = comp + 1;
  // atlas vertices
  struct Key { double x, y; int v; };
  struct KeyHash {
    size_t operator()(const Key& k) const {
      uint64_t a, b;
      memcpy(&a, &k.x, 8), memcpy(&b, &k.y, 8);
      return std::hash<uint64_t>()(a * 0x9E3779B97F4A7C15ULL ^ (b + 0x7F4A7C15ULL) ^ (uint64_t)k.v * 0xBF58476D1CE4E5B9ULL);
    }
  };
  struct KeyEq { bool operator()(const Key& a, const Key& b) const { return a.x == b.x && a.y == b.y && a.v == b.v; } };
  std::unordered_map<Key, int, KeyHash, KeyEq> atlasVertexId;
  std::vector<V2> aVert;
  std::vector<int> aVertMesh, aTri(3 * (size_t)nT);
  for (int t = 0; t < nT; t++)
    for (int k = 0; k < 3; k++) {
      Key key{UV[3 * t + k].x, UV[3 * t + k].y, T[3 * t + k]};
      auto it = atlasVertexId.find(key);
      if (it == atlasVertexId.end()) {
        int id = (int)aVert.size();
        atlasVertexId.emplace(key, id);
        aVert.push_back(UV[3 * t + k]);
        aVertMesh.push_back(T[3 * t + k]);
        aTri[3 * t + k] = id;
      } else aTri[3 * t + k] = it->second;
    }
  // A corner on a texel-centre line makes clipping degenerate; the reference jitters such
  // corners by ~1e-6 texel with rand().  Inputs here must keep corners off those lines.
  for (auto& p : aVert)
    for (int c = 0; c < 2; c++) {
      double s = p[c] * double(c ? H : W) - 0.5;
      if (std::fabs(s - std::round(s)) < 1e-6) { err = "UV corner lies on a texel-centre line (jitter the input UVs)"; return 0; }
    }
  std::unordered_map<uint64_t, int> atlasHalfEdge;
  atlasHalfEdge.reserve(3 * (size_t)nT);
  for (int t = 0; t < nT; t++)
    for (int k = 0; k < 3; k++)
      if (!atlasHalfEdge.emplace(EdgeKey(aTri[3 * t + k], aTri[3 * t + (k + 1) % 3]), 3 * t + k).second) {
        err = "non-oriented manifold atlas mesh";
        return 0;
      }
  const int res = std::max(W, H);
  int lastEdge = 2 * res * res * A.numCharts;
  A.halfEdgeToEdgeIndex.assign(3 * (size_t)nT, -1);
  for (int t = 0; t < nT; t++)
    for (int k = 0; k < 3; k++) {
      int cur = 3 * t + k;
      auto it = atlasHalfEdge.find(EdgeKey(aTri[3 * t + (k + 1) % 3], aTri[3 * t + k]));
      if (it != atlasHalfEdge.end()) {
        if (cur < it->second) A.halfEdgeToEdgeIndex[cur] = A.halfEdgeToEdgeIndex[it->second] = lastEdge++;
      } else A.halfEdgeToEdgeIndex[cur] = lastEdge++;
    }
  // seam / boundary half-edges on the mesh
  std::vector<int> boundaryHalfEdges;
  std::vector<char> isBoundaryHalfEdge(3 * (size_t)nT, 0);
  A.oppositeHalfEdge.assign(3 * (size_t)nT, -1);
  A.isClosed = true;
  for (int t = 0; t < nT; t++)
    for (int k = 0; k < 3; k++) {
      int cur = 3 * t + k;
      auto it = meshHalfEdge.find(EdgeKey(T[3 * t + (k + 1) % 3], T[3 * t + k]));
      if (it != meshHalfEdge.end()) {
        int opp = it->second, o = opp / 3, ok = opp % 3;
        if (cur < opp) A.oppositeHalfEdge[cur] = opp, A.oppositeHalfEdge[opp] = cur;
        bool match = sameUV(3 * t + (k + 1) % 3, 3 * o + ok) && sameUV(3 * t + k, 3 * o + (ok + 1) % 3);
        if (!match && cur < opp) {
          boundaryHalfEdges.push_back(cur);
          boundaryHalfEdges.push_back(opp);
          isBoundaryHalfEdge[cur] = isBoundaryHalfEdge[opp] = 1;
        }
      } else {
        A.isClosed = false;
        A.oppositeHalfEdge[cur] = -1;
        boundaryHalfEdges.push_back(cur);
        isBoundaryHalfEdge[cur] = 1;
      }
    }
  A.boundaryVertexIndex.clear();
  int lastB = 0;
  for (int he : boundaryHalfEdges) {
    int t = he / 3, k = he % 3;
    for (int e = 0; e < 2; e++) {
      int v = T[3 * t + (k + e) % 3];
      if (A.boundaryVertexIndex.find(v) == A.boundaryVertexIndex.end()) A.boundaryVertexIndex[v] = lastB++;
    }
  }
  A.numBoundaryVertices = lastB;
  // per-chart arrays
  A.charts.assign(A.numCharts, AtlasChart());
  for (auto& c : A.charts) c.minCorner = V2(1, 1), c.maxCorner = V2(0, 0);
  std::vector<int> chartVertex(aVert.size(), -1);
  for (int t = 0; t < nT; t++) {
    AtlasChart& C = A.charts[A.triangleChart[t]];
    C.meshTriangleIndices.push_back(t);
    for (int k = 0; k < 3; k++) {
      C.atlasEdgeIndices.push_back(A.halfEdgeToEdgeIndex[3 * t + k]);
      int av = aTri[3 * t + k];
      if (chartVertex[av] == -1) {
        chartVertex[av] = (int)C.vertices.size();
        V2 p = aVert[av];
        for (int c = 0; c < 2; c++) C.minCorner[c] = std::min(p[c], C.minCorner[c]), C.maxCorner[c] = std::max(p[c], C.maxCorner[c]);
        C.vertices.push_back(p);
        C.meshVertexIndices.push_back(aVertMesh[av]);
      }
      C.triangles.push_back(chartVertex[av]);
    }
  }
  for (auto& C : A.charts) {
    V2 mid = (C.minCorner + C.maxCorner) / 2.0;
    mid.x = std::floor(mid.x * W - 0.5), mid.y = std::floor(mid.y * H - 0.5);
    C.originCoords[0] = (int)std::round(mid.x), C.originCoords[1] = (int)std::round(mid.y);
    C.gridOrigin = V2((mid.x + 0.5) / W, (mid.y + 0.5) / H);
    for (int t = 0; t < (int)C.meshTriangleIndices.size(); t++)
      for (int k = 0; k < 3; k++)
        if (isBoundaryHalfEdge[3 * (size_t)C.meshTriangleIndices[t] + k]) C.boundaryHalfEdges.push_back(3 * t + k);
  }
  return 1;
}

// ---------------------------------------------------------------- grid helpers
// GetTriangleIntegerBBox / GetEdgeIntegerBBox (PolygonClipping.h:576-606)
static void IntegerBBox(const V2* p, int n, double icw, double ich, int mn[2], int mx[2]) {
  double fmnx = p[0].x, fmny = p[0].y, fmxx = p[0].x, fmxy = p[0].y;
  for (int i = 1; i < n; i++) fmnx = std::min(fmnx, p[i].x), fmny = std::min(fmny, p[i].y), fmxx = std::max(fmxx, p[i].x), fmxy = std::max(fmxy, p[i].y);
  fmnx = std::max(fmnx, 0.0), fmny = std::max(fmny, 0.0), fmxx = std::min(fmxx, 1.0), fmxy = std::min(fmxy, 1.0);
  mn[0] = (int)std::floor(fmnx * icw), mn[1] = (int)std::floor(fmny * ich);
  mx[0] = (int)std::ceil(fmxx * icw), mx[1] = (int)std::ceil(fmxy * ich);
}

// Inverse of the map (a,b) -> a*(p1-p0) + b*(p2-p0) (GetBaricentricMap, PolygonClipping.h:608-615)
static M2 BarycentricMap(const V2* p) {
  M2 m;
  m(0, 0) = p[1].x - p[0].x, m(0, 1) = p[2].x - p[0].x;
  m(1, 0) = p[1].y - p[0].y, m(1, 1) = p[2].y - p[0].y;
  return m.inv();
}

struct Counters { int texel = 0, interior = 0, deep = 0, boundary = 0, cell = 0, boundaryCell = 0, interiorCell = 0; };

static int FindSegments(const GridChart& g, int j, int i0, int i1, std::vector<Segment>& out) {
  // runs of deep texels in row j, columns [i0, i1] (HierarchyConstruction.inl:412-444)
  out.clear();
  int segStart = -1;
  bool prevDeep = false;
  for (int off = i0; off <= i1; off++) {
    bool cur = g.deep(off, j) != -1;
    if (cur && !prevDeep) segStart = off;
    if ((prevDeep && !cur) || (cur && off == i1)) {
      Segment s;
      s.currentStart = g.texel(segStart, j);
      s.currentEnd = cur ? g.texel(off, j) : g.texel(off - 1, j);
      s.previousStart = g.texel(segStart, j - 1);
      s.nextStart = g.texel(segStart, j + 1);
      s.deepStart = g.deep(segStart, j);
      if (s.previousStart < 0 || s.nextStart < 0) return 0;
      out.push_back(s);
    }
    prevDeep = cur;
  }
  return 1;
}

// InitializeGridChartsActiveNodes (HierarchyConstruction.inl:39-493)
static int BuildGridChart(int chartId, const AtlasChart& ac, GridChart& g, Level& L, Counters& cnt, const BlockInfo& bi,
                          std::string& err) {
  const int w = g.width, h = g.height;
  const double cw = g.cellW, ch = g.cellH;
  g.nodeType.assign((size_t)w * h, -1);
  g.cellType.assign((size_t)(w - 1) * (h - 1), -1);
  auto node = [&](int i, int j) -> int8_t& { return g.nodeType[(size_t)j * w + i]; };
  auto setCell = [&](int i, int j, int8_t v) {
    if (i >= 0 && j >= 0 && i < w - 1 && j < h - 1) g.cellType[(size_t)j * (w - 1) + i] = v;
  };
  const int nTri = (int)ac.triangles.size() / 3;
  // (1) texels covered by a triangle
  for (int t = 0; t < nTri; t++) {
    V2 p[3];
    for (int k = 0; k < 3; k++) p[k] = ac.vertices[ac.triangles[3 * t + k]] - g.corner;
    int mn[2], mx[2];
    IntegerBBox(p, 3, 1.0 / cw, 1.0 / ch, mn, mx);
    M2 bm = BarycentricMap(p);
    for (int j = mn[1]; j <= mx[1]; j++)
      for (int i = mn[0]; i <= mx[0]; i++) {
        if (i < 0 || j < 0 || i >= w || j >= h) continue;
        V2 b = bm * (V2(i * cw, j * ch) - p[0]);
        if (b.x >= 0 && b.y >= 0 && b.x + b.y <= 1) {
          if (node(i, j) != -1) { err = "texel covered twice"; return 0; }
          node(i, j) = 1;
        }
      }
  }
  // (2) texels around cells crossed by boundary edges
  for (int he : ac.boundaryHalfEdges) {
    int t = he / 3, k = he % 3;
    V2 e[2] = {ac.vertices[ac.triangles[3 * t + k]] - g.corner, ac.vertices[ac.triangles[3 * t + (k + 1) % 3]] - g.corner};
    int mn[2], mx[2];
    IntegerBBox(e, 2, 1.0 / cw, 1.0 / ch, mn, mx);
    V2 dir = e[1] - e[0];
    V2 en(dir.y, -dir.x);
    en = en / std::sqrt(dot(en, en));
    double eLevel = (dot(en, e[0]) + dot(en, e[1])) / 2.0;
    for (int c = 0; c < 2; c++)
      for (int j = mn[1]; j <= mx[1]; j++)
        for (int i = mn[0]; i <= mx[0]; i++) {
          V2 n0(i * cw, j * ch), n1 = n0;
          if (c == 0) n1.x += cw;
          else n1.y += ch;
          V2 side = n1 - n0;
          V2 sn(side.y, -side.x);
          sn = sn / std::sqrt(dot(sn, sn));
          double cLevel = (dot(sn, n0) + dot(sn, n1)) / 2.0;
          bool oppEdge = (dot(en, n0) - eLevel) * (dot(en, n1) - eLevel) < 0.0;
          bool oppCell = (dot(sn, e[1]) - cLevel) * (dot(sn, e[0]) - cLevel) < 0.0;
          if (!(oppEdge && oppCell)) continue;
          if (c == 0) setCell(i, j - 1, 0), setCell(i, j, 0);
          else setCell(i - 1, j, 0), setCell(i, j, 0);
          for (int dn = -1; dn <= 1; dn++)
            for (int dc = 0; dc < 2; dc++) {
              int ni[2] = {i, j};
              ni[1 - c] += dn;
              ni[c] += dc;
              ni[0] = std::min(std::max(0, ni[0]), w - 1);
              ni[1] = std::min(std::max(0, ni[1]), h - 1);
              if (node(ni[0], ni[1]) != 1) node(ni[0], ni[1]) = 0;
            }
        }
    if (mn[0] + 1 == mx[0] && mn[1] + 1 == mx[1]) {
      setCell(mn[0], mn[1], 0);
      for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++) {
          int ni = std::min(std::max(0, mn[0] + di), w - 1), nj = std::min(std::max(0, mn[1] + dj), h - 1);
          if (node(ni, nj) != 1) node(ni, nj) = 0;
        }
    }
  }
  // (3) interior cells, (4) deep texels (all four incident cells interior)
  for (int j = 0; j < h - 1; j++)
    for (int i = 0; i < w - 1; i++)
      if (node(i, j) == 1 && node(i + 1, j) == 1 && node(i, j + 1) == 1 && node(i + 1, j + 1) == 1 && g.cell(i, j) == -1)
        g.cellType[(size_t)j * (w - 1) + i] = 1;
  for (int j = 1; j < h - 1; j++)
    for (int i = 1; i < w - 1; i++) {
      bool deep = true;
      for (int di = -1; di < 1; di++)
        for (int dj = -1; dj < 1; dj++)
          if (g.cell(i + di, j + dj) != 1) deep = false;
      if (deep) node(i, j) = 2;
    }
  // (5) raster-order enumeration
  g.texelOffset = cnt.texel, g.interiorTexelOffset = cnt.interior, g.deepOffset = cnt.deep, g.boundaryOffset = cnt.boundary;
  g.cellOffset = cnt.cell;
  g.globalTexel.assign((size_t)w * h, -1);
  g.globalInterior.assign((size_t)w * h, -1);
  g.globalDeep.assign((size_t)w * h, -1);
  g.globalBoundary.assign((size_t)w * h, -1);
  for (int j = 0; j < h; j++)
    for (int i = 0; i < w; i++) {
      size_t q = (size_t)j * w + i;
      int nt = g.nodeType[q];
      if (nt != -1) {
        g.globalTexel[q] = cnt.texel++;
        L.nodeInfo.push_back({chartId, i, j, nt});
      }
      if (nt >= 1) g.globalInterior[q] = cnt.interior++;
      if (nt == 0 || nt == 1) g.globalBoundary[q] = cnt.boundary++;
      if (nt == 2) g.globalDeep[q] = cnt.deep++;
    }
  g.localCell.assign((size_t)(w - 1) * (h - 1), -1);
  g.localInteriorCell.assign((size_t)(w - 1) * (h - 1), -1);
  g.localBoundaryCell.assign((size_t)(w - 1) * (h - 1), -1);
  int nCell = 0, nB = 0, nI = 0;
  for (int j = 0; j < h - 1; j++)
    for (int i = 0; i < w - 1; i++) {
      size_t q = (size_t)j * (w - 1) + i;
      if (g.cellType[q] == -1) continue;
      g.localCell[q] = nCell++;
      std::array<int, 4> gt = {g.texel(i, j), g.texel(i + 1, j), g.texel(i + 1, j + 1), g.texel(i, j + 1)};
      if (gt[0] < 0 || gt[1] < 0 || gt[2] < 0 || gt[3] < 0) { err = "active cell adjacent to inactive texel"; return 0; }
      g.cellIndices.push_back(gt);
      if (g.cellType[q] == 0) {
        g.localBoundaryCell[q] = nB++;
        g.boundaryCellToLocal.push_back(nCell - 1);
      } else {
        g.localInteriorCell[q] = nI++;
        g.interiorCellToLocal.push_back(nCell - 1);
        auto gi = [&](int a, int b) { return g.globalInterior[(size_t)b * w + a]; };
        std::array<int, 4> it = {gi(i, j), gi(i + 1, j), gi(i + 1, j + 1), gi(i, j + 1)};
        if (it[0] < 0 || it[1] < 0 || it[2] < 0 || it[3] < 0) { err = "interior cell adjacent to non-interior texel"; return 0; }
        g.interiorCellCorners.push_back(it);
        g.interiorCellGlobalCorners.push_back(gt);
      }
    }
  cnt.cell += nCell, cnt.boundaryCell += nB, cnt.interiorCell += nI;
  g.numInteriorCells = nI, g.numBoundaryCells = nB;
  // (6) deep raster lines
  for (int j = 0; j < h; j++) {
    bool prev = false;
    int start = -1;
    for (int off = 0; off < w; off++) {
      bool cur = g.node(off, j) == 2;
      if (cur && !prev) start = off;
      if (!cur && prev) {
        Line5 l{g.texel(start, j), g.texel(off - 1, j), g.texel(start, j - 1), g.texel(start, j + 1), g.deep(start, j)};
        if (l.start < 0 || l.end < 0 || l.prev < 0 || l.next < 0) { err = "invalid raster line"; return 0; }
        L.rasterLines.push_back(l);
      }
      prev = cur;
    }
  }
  // (7) band tasks: 16-row bands of 128-wide blocks overlapping by 2 columns
  const int hOff = bi.blockWidth - bi.paddingWidth, vOff = bi.blockHeight - bi.paddingHeight;
  const int nH = ((w - bi.paddingWidth - 1) / hOff) + 1, nV = ((h - bi.paddingHeight - 1) / vOff) + 1;
  BandTasks& T = L.tasks;
  std::vector<Segment> segs;
  for (int bj = 0; bj < nV; bj++) {
    int vS = bj * vOff, vE = std::min((bj + 1) * vOff + bi.paddingHeight + 2 - 1, h - 1);
    int blocksBefore = (int)T.blockBegin.size();
    for (int bI = 0; bI < nH; bI++) {
      int hS = bI * hOff, hE = std::min((bI + 1) * hOff + bi.paddingWidth + 2 - 1, w - 1);
      std::vector<std::vector<Segment>> deepRows, padRows;
      std::vector<int> deepJ, padJ;
      for (int j = vS + 1; j <= vE - 1; j++) {
        if (!FindSegments(g, j, hS + 1, hE - 1, segs)) { err = "invalid block segment"; return 0; }
        if (!segs.empty()) deepRows.push_back(segs), deepJ.push_back(j);
      }
      if (deepRows.empty()) continue;
      for (int j = vS; j <= vE; j++) {
        if (!FindSegments(g, j, hS, hE, segs)) { err = "invalid block segment"; return 0; }
        if (!segs.empty()) padRows.push_back(segs), padJ.push_back(j);
      }
      T.blockBegin.push_back((int)T.linePtr.size() - 1);
      for (size_t k = 0; k < deepRows.size(); k++) {
        T.segments.insert(T.segments.end(), deepRows[k].begin(), deepRows[k].end());
        T.linePtr.push_back((int)T.segments.size());
        T.lineRow.push_back(deepJ[k]);
      }
      T.blockDeepEnd.push_back((int)T.linePtr.size() - 1);
      for (size_t k = 0; k < padRows.size(); k++) {
        T.segments.insert(T.segments.end(), padRows[k].begin(), padRows[k].end());
        T.linePtr.push_back((int)T.segments.size());
        T.lineRow.push_back(padJ[k]);
      }
      T.blockPadEnd.push_back((int)T.linePtr.size() - 1);
    }
    if ((int)T.blockBegin.size() > blocksBefore) T.taskPtr.push_back((int)T.blockBegin.size());
  }
  return 1;
}

// InitializeGridCharts (HierarchyConstruction.inl:496-532) + InitializeBoundaryAndDeepTexelIndexing (:3-35)
static int BuildLevel(const Atlas& A, double cw, double ch, const BlockInfo& bi, Level& L, std::string& err) {
  L.charts.assign(A.charts.size(), GridChart());
  Counters cnt;
  for (size_t c = 0; c < A.charts.size(); c++) {
    const AtlasChart& ac = A.charts[c];
    GridChart& g = L.charts[c];
    int hs[2][2];
    for (int d = 0; d < 2; d++) {
      double cs = d == 0 ? cw : ch;
      hs[d][0] = (int)std::ceil((ac.gridOrigin[d] - ac.minCorner[d]) / cs);
      hs[d][1] = (int)std::ceil((ac.maxCorner[d] - ac.gridOrigin[d]) / cs);
      g.corner[d] = ac.gridOrigin[d] - cs * double(hs[d][0]);
      g.centerOffset[d] = hs[d][0];
      g.cornerCoords[d] = ac.originCoords[d] - hs[d][0];
    }
    g.width = hs[0][0] + hs[0][1] + 1;
    g.height = hs[1][0] + hs[1][1] + 1;
    g.cellW = cw, g.cellH = ch;
    int approx = std::max((int)std::ceil(1.0 / cw), (int)std::ceil(1.0 / ch));
    g.gridIndexOffset = 2 * approx * approx * (int)c;
    if (!BuildGridChart((int)c, ac, g, L, cnt, bi, err)) return 0;
  }
  L.numTexels = cnt.texel, L.numInterior = cnt.interior, L.numDeep = cnt.deep, L.numBoundary = cnt.boundary;
  L.numCells = cnt.cell, L.numBoundaryCells = cnt.boundaryCell, L.numInteriorCells = cnt.interiorCell;
  if (L.numTexels != L.numBoundary + L.numDeep) { err = "boundary and deep texels do not partition the texels"; return 0; }
  L.boundaryAndDeepIndex.assign(L.numTexels, 0);
  L.boundaryGlobalIndex.clear(), L.deepGlobalIndex.clear();
  int g = 0, b = 1, d = -1;
  for (auto& gc : L.charts)
    for (int j = 0; j < gc.height; j++)
      for (int i = 0; i < gc.width; i++) {
        int nt = gc.node(i, j);
        if (nt == 0 || nt == 1) L.boundaryGlobalIndex.push_back(g), L.boundaryAndDeepIndex[g] = b++, g++;
        else if (nt == 2) L.deepGlobalIndex.push_back(g), L.boundaryAndDeepIndex[g] = d--, g++;
      }
  BandSchedule(L.tasks);
  return 1;
}

// InitializeAtlasHierachicalRestriction (ProlongationAndRestriction.inl:266-452)
static int BuildRestriction(const Level& F, Level& C, Csr& boundaryRestriction, std::string& err) {
  const size_t nl = C.rasterLines.size();
  C.restrictionLines.resize(nl);
  C.deepLines.resize(nl);
  for (size_t i = 0; i < nl; i++) {
    const Line5& cl = C.rasterLines[i];
    int len = cl.end - cl.start + 1;
    C.restrictionLines[i].coeff = cl.start;
    C.deepLines[i].coarseStart = cl.coeff;
    C.deepLines[i].coarseEnd = cl.coeff + len - 1;
    const NodeInfo& ni = C.nodeInfo[cl.start];
    const GridChart &fc = F.charts[ni.chart], &cc = C.charts[ni.chart];
    int fi = fc.centerOffset[0] + 2 * (ni.ci - cc.centerOffset[0]);
    int fj = fc.centerOffset[1] + 2 * (ni.cj - cc.centerOffset[1]);
    if (fi - 1 < 0 || fi + 1 > fc.width - 1 || fj - 1 < 0 || fj + 1 > fc.height - 1) { err = "restriction line out of bounds"; return 0; }
    int cur = fc.texel(fi, fj), prv = fc.texel(fi, fj - 1), nxt = fc.texel(fi, fj + 1);
    int curD = fc.deep(fi, fj), prvD = fc.deep(fi, fj - 1), nxtD = fc.deep(fi, fj + 1);
    if (cur < 0 || prv < 0 || nxt < 0 || curD < 0 || prvD < 0 || nxtD < 0) { err = "restriction line on a non-deep fine texel"; return 0; }
    C.restrictionLines[i].start = cur;
    C.restrictionLines[i].end = cur + 2 * (cl.end - cl.start);
    C.restrictionLines[i].prev = prv;
    C.restrictionLines[i].next = nxt;
    C.deepLines[i].fineCurrent = curD, C.deepLines[i].finePrev = prvD, C.deepLines[i].fineNext = nxtD;
  }
  std::vector<Triplet> trip;
  for (size_t k = 0; k < F.charts.size(); k++) {
    const GridChart &fc = F.charts[k], &cc = C.charts[k];
    for (int fj = 0; fj < fc.height; fj++)
      for (int fi = 0; fi < fc.width; fi++) {
        int fg = fc.texel(fi, fj);
        if (fg == -1) continue;
        int _fi = fi - fc.centerOffset[0], _fj = fj - fc.centerOffset[1];
        int si = _fi < 0 ? -1 : (_fi > 0 ? 1 : 0), sj = _fj < 0 ? -1 : (_fj > 0 ? 1 : 0);
        int ci[2], cj[2];
        double wi[2], wj[2];
        if (_fi % 2 == 0) ci[0] = _fi / 2 + cc.centerOffset[0], ci[1] = -1, wi[0] = 1, wi[1] = 0;
        else ci[0] = _fi / 2 + cc.centerOffset[0], ci[1] = (_fi + si) / 2 + cc.centerOffset[0], wi[0] = wi[1] = 0.5;
        if (_fj % 2 == 0) cj[0] = _fj / 2 + cc.centerOffset[1], cj[1] = -1, wj[0] = 1, wj[1] = 0;
        else cj[0] = _fj / 2 + cc.centerOffset[1], cj[1] = (_fj + sj) / 2 + cc.centerOffset[1], wj[0] = wj[1] = 0.5;
        for (int di = 0; di < 2; di++)
          for (int dj = 0; dj < 2; dj++) {
            if (ci[di] == -1 || cj[dj] == -1) continue;
            if (ci[di] < 0 || cj[dj] < 0 || ci[di] >= cc.width || cj[dj] >= cc.height) { err = "coarse texel out of grid"; return 0; }
            int cg = cc.texel(ci[di], cj[dj]);
            if (cg == -1) { err = "coarse texel is inactive"; return 0; }
            int cb = C.boundaryAndDeepIndex[cg] - 1;
            if (cb >= 0) trip.push_back({cb, fg, wi[di] * wj[dj]});
          }
      }
  }
  boundaryRestriction = Csr::FromTriplets(C.numBoundary, F.numTexels, trip);
  return 1;
}

// InitializeAtlasHierachicalProlongation (ProlongationAndRestriction.inl:99-216)
static int BuildProlongation(Level& F, const Level& C, std::string& err) {
  F.prolongationLines.clear();
  for (size_t k = 0; k < F.charts.size(); k++) {
    const GridChart &fc = F.charts[k], &cc = C.charts[k];
    const int w = fc.width;
    for (int j = 0; j < fc.height; j++) {
      bool prev = false;
      int start = -1;
      for (int off = 0; off < w; off++) {
        bool cur = fc.texel(off, j) != -1;
        if (cur && !prev) start = off;
        if ((!cur && prev) || (cur && off == w - 1)) {
          int end = (cur && off == w - 1) ? off + 1 : off;
          ProlongationLine pl;
          pl.start = fc.texel(start, j);
          pl.length = end - start;
          int _fi = start - fc.centerOffset[0], _fj = j - fc.centerOffset[1];
          int _di = (_fi % 2 == 0) ? 0 : 1, _dj = (_fj % 2 == 0) ? 0 : 1;
          int ci = (_fi - _di) / 2 + cc.centerOffset[0], cj = (_fj - _dj) / 2 + cc.centerOffset[1];
          pl.aligned = (_fi % 2 == 0) ? 1 : 0;
          auto ct = [&](int a, int b) { return (a >= 0 && b >= 0 && a < cc.width && b < cc.height) ? cc.texel(a, b) : -1; };
          if (_fj % 2 == 0) {
            pl.prev = pl.next = -1;
            pl.center = ct(ci, cj);
            if (pl.center == -1) { err = "coarse texel inactive (prolongation A)"; return 0; }
          } else {
            pl.center = -1;
            pl.prev = ct(ci, cj);
            pl.next = ct(ci, cj + 1);
            if (pl.prev == -1 || pl.next == -1) { err = "coarse texel inactive (prolongation B/C)"; return 0; }
          }
          F.prolongationLines.push_back(pl);
        }
        prev = cur;
      }
    }
  }
  // every prolongation row must sum to one (ProlongationAndRestriction.inl:173-213)
  std::vector<double> sum(F.numTexels, 0.0);
  for (auto& pl : F.prolongationLines) {
    int off = pl.aligned ? 0 : 1;
    for (int i = 0; i < pl.length; i++) {
      int a = (i + off) / 2, b = (i + off + 1) / 2;
      if (pl.center != -1) {
        if (pl.center + b >= C.numTexels) { err = "prolongation index out of range"; return 0; }
        sum[pl.start + i] += 1.0;
      } else {
        if (pl.prev + b >= C.numTexels || pl.next + b >= C.numTexels) { err = "prolongation index out of range"; return 0; }
        sum[pl.start + i] += 1.0;
      }
      (void)a;
    }
  }
  for (int i = 0; i < F.numTexels; i++)
    if (std::fabs(sum[i] - 1.0) > 1e-10) { err = "prolongation rows do not sum to one"; return 0; }
  return 1;
}

int BuildHierarchy(const TexturedMesh& mesh, int W, int H, int numLevels, const BlockInfo& bi, Hierarchy& h, std::string& err) {
  if (numLevels < 1) { err = "need at least one level"; return 0; }
  h.width = W, h.height = H;
  if (!BuildAtlas(mesh, W, H, h.atlas, err)) return 0;
  h.levels.assign(numLevels, Level());
  for (int l = 0; l < numLevels; l++) {
    double f = double(1 << l);
    if (!BuildLevel(h.atlas, f / W, f / H, bi, h.levels[l], err)) { err = "level " + std::to_string(l) + ": " + err; return 0; }
  }
  for (int l = 0; l + 1 < numLevels; l++) {
    if (!BuildRestriction(h.levels[l], h.levels[l + 1], h.levels[l].boundaryRestriction, err)) return 0;
    if (!BuildProlongation(h.levels[l], h.levels[l + 1], err)) return 0;
  }
  return 1;
}

// Pipeline schedule of the band tasks (the GPU band kernel): block b of a task runs its local steps
// t = 0 .. max(nDeep + 4, nPad + 6) - 1 at task steps blockStart[b] + t.  At local step t it touches (reads
// or writes) chart rows Y - 1 .. Y + 1 of deep row i = t - 2w (passes w = 0, 1, 2) and of padded row
// t - 6 (residual).  A block starts once every row it shares with the block before it is touched
// by it for the last time, so that each row's accesses keep the reference's block order (blocks
// further apart share no columns), and once the wave group it takes (every kBandGroups-th block)
// is free.  taskSteps[t]: task steps of task t.
void BandSchedule(BandTasks& T) {
  // row elements: every line's segments expanded texel by texel
  const int nl = (int)T.linePtr.size() - 1;
  T.lineElem.assign(1, 0), T.elems.clear();
  for (int l = 0; l < nl; l++) {
    for (int s = T.linePtr[l]; s < T.linePtr[l + 1]; s++) {
      const Segment& sg = T.segments[s];
      for (int o = 0; o <= sg.currentEnd - sg.currentStart; o++)
        T.elems.insert(T.elems.end(), {sg.currentStart + o, sg.previousStart + o, sg.nextStart + o, o == 0 ? -(sg.deepStart + o) - 1 : sg.deepStart + o});
    }
    T.lineElem.push_back((int)T.elems.size() / 4);
  }
  const int nb = (int)T.blockBegin.size(), nt = (int)T.taskPtr.size() - 1;
  std::vector<int>& start = T.blockStart;
  std::vector<int>& steps = T.taskSteps;
  start.assign(nb, 0), steps.assign(std::max(nt, 0), 0);
  auto blockSteps = [&](int b) {
    return std::max(T.blockDeepEnd[b] - T.blockBegin[b] + 4, T.blockPadEnd[b] - T.blockDeepEnd[b] + 6);
  };
  // first and last local step at which block b touches each chart row (rows as a span [y0, y0 + n))
  struct Touch {
    int y0 = 0;
    std::vector<int> first, last;
  };
  auto touches = [&](int b) {
    Touch tc;
    int lo = INT32_MAX, hi = INT32_MIN;
    for (int l = T.blockBegin[b]; l < T.blockPadEnd[b]; l++) lo = std::min(lo, T.lineRow[l]), hi = std::max(hi, T.lineRow[l]);
    tc.y0 = lo - 1;
    tc.first.assign(hi - lo + 3, INT32_MAX), tc.last.assign(hi - lo + 3, -1);
    auto mark = [&](int y, int t) {
      for (int dy = -1; dy <= 1; dy++) {
        const int k = y + dy - tc.y0;
        tc.first[k] = std::min(tc.first[k], t), tc.last[k] = std::max(tc.last[k], t);
      }
    };
    for (int i = 0; i < T.blockDeepEnd[b] - T.blockBegin[b]; i++)
      for (int w = 0; w < 3; w++) mark(T.lineRow[T.blockBegin[b] + i], i + 2 * w);
    for (int i = 0; i < T.blockPadEnd[b] - T.blockDeepEnd[b]; i++) mark(T.lineRow[T.blockDeepEnd[b] + i], i + 6);
    return tc;
  };
  for (int t = 0; t < nt; t++) {
    Touch prev;
    for (int b = T.taskPtr[t]; b < T.taskPtr[t + 1]; b++) {
      Touch cur = touches(b);
      int s0 = 0;
      if (b > T.taskPtr[t]) {
        s0 = start[b - 1];
        for (size_t k = 0; k < cur.first.size(); k++) {
          const int y = cur.y0 + (int)k, kp = y - prev.y0;
          if (cur.last[k] < 0 || kp < 0 || kp >= (int)prev.last.size() || prev.last[kp] < 0) continue;
          s0 = std::max(s0, start[b - 1] + prev.last[kp] - cur.first[k] + 1);
        }
        if (b - kBandGroups >= T.taskPtr[t]) s0 = std::max(s0, start[b - kBandGroups] + blockSteps(b - kBandGroups));
      }
      start[b] = s0;
      steps[t] = std::max(steps[t], s0 + blockSteps(b));
      prev = std::move(cur);
    }
  }
}

}  // namespace gdtp
