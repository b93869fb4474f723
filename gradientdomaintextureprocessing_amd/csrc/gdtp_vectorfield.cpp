// Normalised-gradient integration samples (VectorFieldIntergration.inl:71-438).
//
// Interior cells: one sample per (mesh triangle ∩ cell) piece at the piece's area barycentre, with
// the integrated texel-basis gradients of the piece (the 6-point rule is exact: the gradients are
// linear), as the reference builds them (:114-214).  Boundary cells: the reference samples its
// quadratic boundary elements (:215-420); away from seams those carry the bilinear field, so the
// piece is sampled like an interior cell; on pieces touching a seam each quadratic triangle of
// the piece is one sample at its barycentre, its node values expanded into texel weights.
#include "gdtp_vectorfield.h"

#include <omp.h>

#include <algorithm>
#include <cmath>

namespace gdtp {

namespace {

const double kQP[6][2] = {{0.816847572980459, 0.0915762135097705}, {0.0915762135097705, 0.816847572980459},
                          {0.0915762135097705, 0.0915762135097705}, {0.108103018168070, 0.445948490915965},
                          {0.445948490915965, 0.108103018168070},  {0.445948490915965, 0.445948490915965}};
const double kQW[6] = {0.109951743655322, 0.109951743655322, 0.109951743655322,
                       0.223381589678011, 0.223381589678011, 0.223381589678011};

inline V2 Q1Grad(int k, double x, double y) {
  switch (k) {
    case 0: return V2(y - 1, x - 1);
    case 1: return V2(1 - y, -x);
    case 2: return V2(y, x);
    default: return V2(-y, 1 - x);
  }
}

struct ChartSamples {
  std::vector<int> cellCorners, genPtr{0}, genTexel;
  std::vector<float> cellData, genTensor, genG, genV;
};

void CellSample(const std::vector<V2>& poly, int i, int j, double cw, double ch, const M2& G, const std::array<int, 4>& corners,
                ChartSamples& out) {
  const M2 Gi = G.inv();
  const double gscale = std::sqrt(G.det());
  V2 bary(0, 0);
  double area = 0;
  V2 v[4];
  for (size_t p = 2; p < poly.size(); p++) {
    const V2 d0 = poly[p - 1] - poly[0], d1 = poly[p] - poly[0];
    const double a = std::fabs(cross(d0, d1)) / 2.0;
    bary = bary + (poly[0] + poly[p - 1] + poly[p]) * (a / 3.0);
    area += a;
    for (int s = 0; s < 6; s++) {
      const V2 q = poly[0] + d0 * kQP[s][0] + d1 * kQP[s][1];
      const double x = q.x / cw - i, y = q.y / ch - j;
      for (int k = 0; k < 4; k++) {
        const V2 gr = Q1Grad(k, x, y);
        v[k] = v[k] + V2(gr.x / cw, gr.y / ch) * (kQW[s] * a * gscale);
      }
    }
  }
  bary = bary / area;
  const double px = bary.x / cw - i, py = bary.y / ch - j;
  for (int k = 0; k < 4; k++) out.cellCorners.push_back(corners[k]);
  const float d[5] = {(float)px, (float)py, (float)Gi(0, 0), (float)(0.5 * (Gi(0, 1) + Gi(1, 0))), (float)Gi(1, 1)};
  out.cellData.insert(out.cellData.end(), d, d + 5);
  for (int k = 0; k < 4; k++) out.cellData.push_back((float)v[k].x), out.cellData.push_back((float)v[k].y);
}

void GeneralSample(const SeamElement& el, const M2& G, const std::vector<int>& boundaryGlobal, ChartSamples& out) {
  const V2 e1 = el.p[1] - el.p[0], e2 = el.p[2] - el.p[0];
  const double det = cross(e1, e2);
  const V2 g1(e2.y / det, -e2.x / det), g2(-e1.y / det, e1.x / det), g0(-g1.x - g2.x, -g1.y - g2.y);
  // quadratic basis gradients at the barycentre (also their means over the triangle: they are linear)
  const V2 grad[6] = {g0 * (1.0 / 3), g1 * (1.0 / 3), g2 * (1.0 / 3), (g0 + g1) * (4.0 / 3), (g1 + g2) * (4.0 / 3), (g2 + g0) * (4.0 / 3)};
  const double scale = std::fabs(det) / 2.0 * std::sqrt(G.det());
  std::vector<std::pair<int, std::pair<V2, V2>>> acc;
  for (int n = 0; n < 6; n++)
    for (const auto& w : el.w[n]) {
      const int t = boundaryGlobal[w.first];
      auto it = std::find_if(acc.begin(), acc.end(), [&](const auto& a) { return a.first == t; });
      if (it == acc.end()) acc.push_back({t, {V2(0, 0), V2(0, 0)}}), it = acc.end() - 1;
      it->second.first = it->second.first + grad[n] * w.second;
      it->second.second = it->second.second + grad[n] * (w.second * scale);
    }
  const M2 Gi = G.inv();
  out.genTensor.push_back((float)Gi(0, 0)), out.genTensor.push_back((float)(0.5 * (Gi(0, 1) + Gi(1, 0)))), out.genTensor.push_back((float)Gi(1, 1));
  for (const auto& a : acc) {
    out.genTexel.push_back(a.first);
    out.genG.push_back((float)a.second.first.x), out.genG.push_back((float)a.second.first.y);
    out.genV.push_back((float)a.second.second.x), out.genV.push_back((float)a.second.second.y);
  }
  out.genPtr.push_back((int)out.genTexel.size());
}

}  // namespace

int BuildVectorFieldSamples(const Hierarchy& h, const std::vector<std::vector<M2>>& metric, VectorFieldSamples& out,
                            std::string& err) {
  const Level& L0 = h.levels[0];
  const int nCharts = (int)L0.charts.size();
  SeamContext seams;
  if (!BuildSeamContext(h, seams, err)) return 0;
  std::vector<ChartSamples> cs(nCharts);
  std::vector<std::string> errs(nCharts);
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nCharts; c++) {
    const GridChart& g = L0.charts[c];
    const AtlasChart& ac = h.atlas.charts[c];
    const int nTri = (int)ac.triangles.size() / 3;
    std::vector<V2> poly;
    std::vector<SeamElement> els;
    for (int t = 0; t < nTri && errs[c].empty(); t++) {
      V2 p[3];
      for (int k = 0; k < 3; k++) p[k] = ac.vertices[ac.triangles[3 * t + k]] - g.corner;
      const M2& G = metric[c][t];
      const double fmn[2] = {std::min({p[0].x, p[1].x, p[2].x}), std::min({p[0].y, p[1].y, p[2].y})};
      const double fmx[2] = {std::max({p[0].x, p[1].x, p[2].x}), std::max({p[0].y, p[1].y, p[2].y})};
      const int mn0 = (int)std::floor(std::max(fmn[0], 0.0) / g.cellW), mn1 = (int)std::floor(std::max(fmn[1], 0.0) / g.cellH);
      const int mx0 = (int)std::ceil(std::min(fmx[0], 1.0) / g.cellW), mx1 = (int)std::ceil(std::min(fmx[1], 1.0) / g.cellH);
      for (int j = mn1; j < mx1 && errs[c].empty(); j++)
        for (int i = mn0; i < mx0; i++) {
          if (i < 0 || j < 0 || i >= g.width - 1 || j >= g.height - 1) continue;
          const size_t q = (size_t)j * (g.width - 1) + i;
          const int li = g.localInteriorCell[q], lb = g.localBoundaryCell[q];
          if (li < 0 && lb < 0) continue;
          poly.assign(p, p + 3);
          if (!ClipToCell(poly, i, j, g.cellW, g.cellH)) continue;
          if (li >= 0) {
            CellSample(poly, i, j, g.cellW, g.cellH, G, g.interiorCellGlobalCorners[li], cs[c]);
            continue;
          }
          els.clear();
          const int rc = SeamPieceElements(seams, c, t, i, j, poly, els, errs[c]);
          if (rc < 0) break;
          if (rc == 0) CellSample(poly, i, j, g.cellW, g.cellH, G, g.cellIndices[g.localCell[q]], cs[c]);
          else
            for (const SeamElement& el : els) GeneralSample(el, G, L0.boundaryGlobalIndex, cs[c]);
        }
    }
  }
  for (int c = 0; c < nCharts; c++)
    if (!errs[c].empty()) { err = errs[c]; return 0; }
  out = VectorFieldSamples();
  out.invCellW = 1.0 / L0.charts[0].cellW, out.invCellH = 1.0 / L0.charts[0].cellH;
  for (int c = 0; c < nCharts; c++) {
    const ChartSamples& s = cs[c];
    out.cellCorners.insert(out.cellCorners.end(), s.cellCorners.begin(), s.cellCorners.end());
    out.cellData.insert(out.cellData.end(), s.cellData.begin(), s.cellData.end());
    const int base = (int)out.genTexel.size();
    for (size_t k = 1; k < s.genPtr.size(); k++) out.genPtr.push_back(base + s.genPtr[k]);
    out.genTexel.insert(out.genTexel.end(), s.genTexel.begin(), s.genTexel.end());
    out.genTensor.insert(out.genTensor.end(), s.genTensor.begin(), s.genTensor.end());
    out.genG.insert(out.genG.end(), s.genG.begin(), s.genG.end());
    out.genV.insert(out.genV.end(), s.genV.begin(), s.genV.end());
  }
  return 1;
}

}  // namespace gdtp
