// Mass / stiffness assembly over atlas-grid cells clipped to mesh triangles.
#include "gdtp_assembly.h"

#include <omp.h>

#include "gdtp_boundary.h"

namespace gdtp {

// Degree-4, six-point rule on the unit right triangle (weights sum to one); the same classical
// symmetric rule the reference uses (RightTriangleQuadrature.h:6-23).
static const double kQuadPos[6][2] = {{0.816847572980459, 0.0915762135097705}, {0.0915762135097705, 0.816847572980459},
                                      {0.0915762135097705, 0.0915762135097705}, {0.108103018168070, 0.445948490915965},
                                      {0.445948490915965, 0.108103018168070},  {0.445948490915965, 0.445948490915965}};
static const double kQuadW[6] = {0.109951743655322, 0.109951743655322, 0.109951743655322,
                                 0.223381589678011, 0.223381589678011, 0.223381589678011};

// Bilinear texel basis on the unit cell; corner order (0,0),(1,0),(1,1),(0,1) (Basis.h:165-206).
static inline double Q1(int k, double x, double y) {
  switch (k) {
    case 0: return (1 - x) * (1 - y);
    case 1: return x * (1 - y);
    case 2: return x * y;
    default: return (1 - x) * y;
  }
}
static inline V2 Q1Grad(int k, double x, double y) {
  switch (k) {
    case 0: return V2(y - 1, x - 1);
    case 1: return V2(1 - y, -x);
    case 2: return V2(y, x);
    default: return V2(-y, 1 - x);
  }
}

void ComputeParameterMetricFromSurface(const TexturedMesh& mesh, const Atlas& atlas, const std::vector<M2>& g,
                                       std::vector<std::vector<M2>>& metric) {
  metric.assign(atlas.charts.size(), {});
  for (size_t c = 0; c < atlas.charts.size(); c++) {
    const AtlasChart& ch = atlas.charts[c];
    metric[c].resize(ch.meshTriangleIndices.size());
    for (size_t k = 0; k < ch.meshTriangleIndices.size(); k++) {
      int t = ch.meshTriangleIndices[k];
      V2 d0 = mesh.uv[3 * t + 1] - mesh.uv[3 * t], d1 = mesh.uv[3 * t + 2] - mesh.uv[3 * t];
      M2 D;  // columns are the UV edge vectors
      D(0, 0) = d0.x, D(1, 0) = d0.y, D(0, 1) = d1.x, D(1, 1) = d1.y;
      M2 Di = D.inv();
      metric[c][k] = Di.T() * g[t] * Di;
    }
  }
}

void ComputeParameterMetric(const TexturedMesh& mesh, const Atlas& atlas, int mode, std::vector<std::vector<M2>>& metric) {
  const int nT = mesh.numTriangles();
  std::vector<M2> g(nT);
  double total = 0;
  for (int t = 0; t < nT; t++) {
    M2 m;
    if (mode == METRIC_UNIFORM) {
      V2 d[2] = {mesh.uv[3 * t + 1] - mesh.uv[3 * t], mesh.uv[3 * t + 2] - mesh.uv[3 * t]};
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) m(a, b) = dot(d[a], d[b]);
    } else {
      const V3 &p0 = mesh.vertices[mesh.triangles[3 * t]], &p1 = mesh.vertices[mesh.triangles[3 * t + 1]],
               &p2 = mesh.vertices[mesh.triangles[3 * t + 2]];
      V3 d[2] = {p1 - p0, p2 - p0};
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) m(a, b) = dot(d[a], d[b]);
    }
    total += std::sqrt(m.det()) / 2;
    g[t] = m;
  }
  for (auto& m : g)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) m(a, b) /= total;
  ComputeParameterMetricFromSurface(mesh, atlas, g, metric);
}

// Sutherland-Hodgman clip of a convex polygon against {p : dot(p, n) > offset}.
void ClipConvex(std::vector<V2>& poly, V2 n, double offset) {
  std::vector<V2> out;
  if (poly.empty()) return;
  V2 prev = poly.back();
  double pl = dot(prev, n) - offset;
  bool pin = pl > 0;
  for (const V2& cur : poly) {
    double cl = dot(cur, n) - offset;
    bool cin = cl > 0;
    if (cin != pin) {
      double a = -cl / (pl - cl);
      out.push_back(prev * a + cur * (1.0 - a));
    }
    if (cin) out.push_back(cur);
    prev = cur, pl = cl, pin = cin;
  }
  poly.swap(out);
}

// Clip to the primal cell [i cw,(i+1) cw] x [j ch,(j+1) ch] (ClipTriangleToPrimalCell, PolygonClipping.h:85-118).
bool ClipToCell(std::vector<V2>& poly, int i, int j, double cw, double ch) {
  ClipConvex(poly, V2(0, 1), j * ch);
  if (poly.empty()) return false;
  ClipConvex(poly, V2(0, -1), -(j + 1) * ch);
  if (poly.empty()) return false;
  ClipConvex(poly, V2(1, 0), i * cw);
  if (poly.empty()) return false;
  ClipConvex(poly, V2(-1, 0), -(i + 1) * cw);
  return !poly.empty();
}

// Integrate the 4x4 bilinear mass / stiffness over a convex polygon inside cell (i,j).
static void IntegrateCellPolygon(const std::vector<V2>& poly, int i, int j, double cw, double ch, const M2& Ginv, double gscale,
                                 double mass[4][4], double stiff[4][4]) {
  for (size_t p = 2; p < poly.size(); p++) {
    V2 d0 = poly[p - 1] - poly[0], d1 = poly[p] - poly[0];
    double area = std::fabs(cross(d0, d1)) * gscale / 2.0;
    double val[4][6];
    V2 grd[4][6];
    for (int s = 0; s < 6; s++) {
      V2 q = poly[0] + d0 * kQuadPos[s][0] + d1 * kQuadPos[s][1];
      double x = q.x / cw - i, y = q.y / ch - j;
      for (int k = 0; k < 4; k++) {
        val[k][s] = Q1(k, x, y);
        V2 g = Q1Grad(k, x, y);
        grd[k][s] = V2(g.x / cw, g.y / ch);
      }
    }
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < 4; l++) {
        double m = 0, st = 0;
        for (int s = 0; s < 6; s++) {
          m += val[k][s] * val[l][s] * kQuadW[s];
          st += dot(grd[l][s], Ginv * grd[k][s]) * kQuadW[s];
        }
        mass[l][k] += m * area;
        stiff[l][k] += st * area;
      }
  }
}

namespace {
// boundary-row entries of the cell-based stiffness products of one chart: (row, cell) + K values
struct CellRows {
  std::vector<int> row, cell;
  std::vector<double> val;  // K per entry
};
}  // namespace

int AssembleFineOperator(const TexturedMesh& mesh, const Hierarchy& h, const std::vector<std::vector<M2>>& metric, FineOperator& op,
                         std::string& err, const double* signal, int K, CellStiffnessRhs* cellRhs, bool deviceInterior) {
  const Level& L0 = h.levels[0];
  const int nCharts = (int)L0.charts.size();
  const long N = L0.numTexels;
  const bool cells = cellRhs && signal && K > 0;
  if (cells && deviceInterior) {
    err = "the cell-based stiffness is assembled on the host";
    return 0;
  }
  // device path: interior cells (and with them every deep row) on the GPU; the host keeps the
  // boundary cells and needs the matrices of the interior cells that touch a boundary texel
  std::vector<int> chartCellBase(nCharts + 1, 0), wantedPtr(nCharts + 1, 0), wanted;
  std::vector<double> wantedM, wantedS;
  if (deviceInterior) {
    for (int c = 0; c < nCharts; c++) {
      const GridChart& g = L0.charts[c];
      chartCellBase[c + 1] = chartCellBase[c] + g.numInteriorCells;
      for (int ic = 0; ic < g.numInteriorCells; ic++) {
        bool touches = false;
        for (int k = 0; k < 4; k++) touches = touches || L0.boundaryAndDeepIndex[g.interiorCellGlobalCorners[ic][k]] > 0;
        if (touches) wanted.push_back(chartCellBase[c] + ic);
      }
      wantedPtr[c + 1] = (int)wanted.size();
    }
    AssembleInteriorCellsOnDevice(h, metric, wanted, op.deepMass, op.deepStiffness, wantedM, wantedS);
  }
  std::vector<CellRows> crows(cells ? nCharts : 0);
  if (cells) {
    cellRhs->K = K, cellRhs->numCells = L0.numCells;
    cellRhs->deep.assign((size_t)K * 4 * L0.numDeep, 0.0);
  }
  // S_c f_k at corner k of a cell whose corners (order (0,0),(1,0),(1,1),(0,1)) are texels gl
  auto cellProducts = [&](const std::array<double, 16>& S, const std::array<int, 4>& gl, int k, double* out) {
    for (int kk = 0; kk < K; kk++) {
      double p = 0;
      for (int l = 0; l < 4; l++) p += S[4 * k + l] * signal[kk * N + gl[l]];
      out[kk] = p;
    }
  };
  if (!deviceInterior) {
    op.deepMass.assign(10 * (size_t)L0.numDeep, 0.0);
    op.deepStiffness.assign(10 * (size_t)L0.numDeep, 0.0);
  }
  std::vector<std::vector<Triplet>> bbM(nCharts), bbS(nCharts), bdM(nCharts), bdS(nCharts);
  std::vector<std::string> errs(nCharts);
  SeamContext seams;
  if (!BuildSeamContext(h, seams, err)) return 0;
  const int offI[4] = {0, 1, 1, 0}, offJ[4] = {0, 0, 1, 1};
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nCharts; c++) {
    const GridChart& g = L0.charts[c];
    const AtlasChart& ac = h.atlas.charts[c];
    const int nTri = (int)ac.triangles.size() / 3;
    const int nHostCells = deviceInterior ? 0 : g.numInteriorCells;  // interior cells integrated here
    std::vector<std::array<double, 16>> cellM(nHostCells, std::array<double, 16>{}), cellS(nHostCells, std::array<double, 16>{});
    std::vector<std::array<double, 16>> bcellM(g.numBoundaryCells, std::array<double, 16>{}), bcellS(g.numBoundaryCells, std::array<double, 16>{});
    std::vector<V2> poly;
    for (int t = 0; t < nTri && errs[c].empty(); t++) {
      V2 p[3];
      for (int k = 0; k < 3; k++) p[k] = ac.vertices[ac.triangles[3 * t + k]] - g.corner;
      const M2& G = metric[c][t];
      M2 Ginv = G.inv();
      double gscale = std::sqrt(G.det());
      double fmn[2] = {std::min({p[0].x, p[1].x, p[2].x}), std::min({p[0].y, p[1].y, p[2].y})};
      double fmx[2] = {std::max({p[0].x, p[1].x, p[2].x}), std::max({p[0].y, p[1].y, p[2].y})};
      int mn0 = (int)std::floor(std::max(fmn[0], 0.0) / g.cellW), mn1 = (int)std::floor(std::max(fmn[1], 0.0) / g.cellH);
      int mx0 = (int)std::ceil(std::min(fmx[0], 1.0) / g.cellW), mx1 = (int)std::ceil(std::min(fmx[1], 1.0) / g.cellH);
      for (int j = mn1; j < mx1 && errs[c].empty(); j++)
        for (int i = mn0; i < mx0; i++) {
          if (i < 0 || j < 0 || i >= g.width - 1 || j >= g.height - 1) continue;
          size_t q = (size_t)j * (g.width - 1) + i;
          int li = g.localInteriorCell[q], lb = g.localBoundaryCell[q];
          if (li < 0 && lb < 0) continue;
          if (li >= 0 && deviceInterior) continue;
          poly.assign(p, p + 3);
          if (!ClipToCell(poly, i, j, g.cellW, g.cellH)) continue;
          if (lb >= 0) {  // boundary cell: pieces touching a seam carry the averaged seam values
            const size_t n0 = bbS[c].size();
            int rc = AssembleSeamPiece(seams, c, t, i, j, poly, G, bbM[c], bbS[c], errs[c]);
            if (rc < 0) break;
            if (rc > 0) {
              if (cells) {  // the piece's texel-space stiffness applied to the signal, charged to this cell
                const int cell = g.cellOffset + g.localCell[q];
                for (size_t e = n0; e < bbS[c].size(); e++) {
                  const Triplet& tr = bbS[c][e];
                  const long col = L0.boundaryGlobalIndex[tr.c];
                  crows[c].row.push_back(tr.r), crows[c].cell.push_back(cell);
                  for (int kk = 0; kk < K; kk++) crows[c].val.push_back(tr.v * signal[kk * N + col]);
                }
              }
              continue;
            }
          }
          double m[4][4] = {}, s[4][4] = {};
          IntegrateCellPolygon(poly, i, j, g.cellW, g.cellH, Ginv, gscale, m, s);
          auto& CM = li >= 0 ? cellM[li] : bcellM[lb];
          auto& CS = li >= 0 ? cellS[li] : bcellS[lb];
          for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) CM[4 * a + b] += m[a][b], CS[4 * a + b] += s[a][b];
        }
    }
    if (!errs[c].empty()) continue;
    // interior cells: deep rows and the boundary rows they touch (MassAndStiffness.h:277-353)
    auto boundaryRow = [&](int b, const std::array<int, 4>& gl, int k, const double* CM, const double* CS) {
      for (int l = 0; l < 4; l++) {
        int nbd = L0.boundaryAndDeepIndex[gl[l]];
        if (nbd < 0) {
          bdM[c].push_back({b, gl[l], CM[4 * k + l]});
          bdS[c].push_back({b, gl[l], CS[4 * k + l]});
        } else {
          bbM[c].push_back({b, nbd - 1, CM[4 * k + l]});
          bbS[c].push_back({b, nbd - 1, CS[4 * k + l]});
        }
      }
    };
    if (deviceInterior)  // deep rows came from the device; boundary rows of the cells that touch them
      for (int w = wantedPtr[c]; w < wantedPtr[c + 1]; w++) {
        const auto& gl = g.interiorCellGlobalCorners[wanted[w] - chartCellBase[c]];
        for (int k = 0; k < 4; k++) {
          const int bd = L0.boundaryAndDeepIndex[gl[k]];
          if (bd > 0) boundaryRow(bd - 1, gl, k, &wantedM[16 * (size_t)w], &wantedS[16 * (size_t)w]);
        }
      }
    std::vector<double> prod(std::max(K, 1));
    for (int ic = 0; ic < nHostCells; ic++) {
      const auto& gl = g.interiorCellGlobalCorners[ic];
      const int cell = g.cellOffset + g.interiorCellToLocal[ic];
      for (int k = 0; k < 4; k++) {
        int bd = L0.boundaryAndDeepIndex[gl[k]];
        if (cells) {
          cellProducts(cellS[ic], gl, k, prod.data());
          if (bd < 0) {
            const int slot = (1 - offJ[k]) * 2 + (1 - offI[k]);  // where this cell sits around texel gl[k]
            for (int kk = 0; kk < K; kk++) cellRhs->deep[(size_t)kk * 4 * L0.numDeep + 4 * (size_t)(-bd - 1) + slot] = prod[kk];
          } else {
            crows[c].row.push_back(bd - 1), crows[c].cell.push_back(cell);
            crows[c].val.insert(crows[c].val.end(), prod.begin(), prod.begin() + K);
          }
        }
        if (bd < 0) {
          int d = -bd - 1;
          for (int l = 0; l < 4; l++) {
            int slot = (offJ[l] - offJ[k] + 1) * 3 + (offI[l] - offI[k] + 1);
            op.deepMass[10 * (size_t)d + slot] += cellM[ic][4 * k + l];
            op.deepStiffness[10 * (size_t)d + slot] += cellS[ic][4 * k + l];
          }
        } else {
          boundaryRow(bd - 1, gl, k, cellM[ic].data(), cellS[ic].data());
        }
      }
    }
    // boundary cells (pieces away from seams)
    AssembleBoundaryCells(h, c, bcellM, bcellS, bbM[c], bbS[c], errs[c]);
    if (cells)
      for (int bc = 0; bc < g.numBoundaryCells; bc++) {
        const int local = g.boundaryCellToLocal[bc];
        const auto& gl = g.cellIndices[local];
        for (int k = 0; k < 4; k++) {
          cellProducts(bcellS[bc], gl, k, prod.data());
          crows[c].row.push_back(L0.boundaryAndDeepIndex[gl[k]] - 1), crows[c].cell.push_back(g.cellOffset + local);
          crows[c].val.insert(crows[c].val.end(), prod.begin(), prod.begin() + K);
        }
      }
  }
  for (int c = 0; c < nCharts; c++)
    if (!errs[c].empty()) { err = errs[c]; return 0; }
  auto merge = [&](std::vector<std::vector<Triplet>>& v) {
    std::vector<Triplet> all;
    for (auto& x : v) all.insert(all.end(), x.begin(), x.end());
    return all;
  };
  if (cells) {
    // deep texels: the cells (i-1,j-1) and (i-1,j) (their right neighbours follow in the same row)
    cellRhs->deepCells.assign(2 * (size_t)L0.numDeep, -1);
    for (const GridChart& g : L0.charts)
      for (int j = 1; j + 1 < g.height; j++)
        for (int i = 1; i + 1 < g.width; i++) {
          const int d = g.deep(i, j);
          if (d < 0) continue;
          cellRhs->deepCells[2 * (size_t)d] = g.cellOffset + g.localCell[(size_t)(j - 1) * (g.width - 1) + i - 1];
          cellRhs->deepCells[2 * (size_t)d + 1] = g.cellOffset + g.localCell[(size_t)j * (g.width - 1) + i - 1];
        }
    // boundary rows: entries sorted by (row, cell), duplicates (several pieces of a cell) summed
    std::vector<std::array<int, 3>> ent;  // row, cell, (chart, entry) packed below
    std::vector<std::pair<int, size_t>> src;
    for (int c = 0; c < nCharts; c++)
      for (size_t e = 0; e < crows[c].row.size(); e++) ent.push_back({crows[c].row[e], crows[c].cell[e], (int)src.size()}), src.push_back({c, e});
    std::sort(ent.begin(), ent.end(), [](const auto& a, const auto& b) { return a[0] != b[0] ? a[0] < b[0] : a[1] < b[1]; });
    std::vector<int> rowOf;
    std::vector<double> acc;  // K per merged entry
    for (size_t e = 0; e < ent.size(); e++) {
      const bool same = !rowOf.empty() && rowOf.back() == ent[e][0] && cellRhs->bCell.back() == ent[e][1];
      if (!same) {
        rowOf.push_back(ent[e][0]), cellRhs->bCell.push_back(ent[e][1]);
        acc.resize(acc.size() + K, 0.0);
      }
      const auto& sr = src[ent[e][2]];
      const double* v = &crows[sr.first].val[sr.second * K];
      double* a = &acc[acc.size() - K];
      for (int kk = 0; kk < K; kk++) a[kk] += v[kk];
    }
    const size_t nnz = rowOf.size();
    cellRhs->bRowptr.assign(L0.numBoundary + 1, 0);
    for (int r : rowOf) cellRhs->bRowptr[r + 1]++;
    for (int r = 0; r < L0.numBoundary; r++) cellRhs->bRowptr[r + 1] += cellRhs->bRowptr[r];
    cellRhs->bVal.assign((size_t)K * nnz, 0.0);
    for (size_t e = 0; e < nnz; e++)
      for (int kk = 0; kk < K; kk++) cellRhs->bVal[(size_t)kk * nnz + e] = acc[e * K + kk];
  }
  {
    auto t = merge(bbM); op.bbMass = Csr::FromTriplets(L0.numBoundary, L0.numBoundary, t);
    t = merge(bbS); op.bbStiffness = Csr::FromTriplets(L0.numBoundary, L0.numBoundary, t);
    t = merge(bdM); op.bdMass = Csr::FromTriplets(L0.numBoundary, L0.numTexels, t);
    t = merge(bdS); op.bdStiffness = Csr::FromTriplets(L0.numBoundary, L0.numTexels, t);
  }
  return 1;
}

}  // namespace gdtp
