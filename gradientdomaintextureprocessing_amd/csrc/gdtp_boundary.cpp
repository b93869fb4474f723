// Boundary-cell assembly and seam coupling (see gdtp_boundary.h for the model).
//
// Reference construction this follows (BoundaryQuadraticElements.inl): seam nodes are the
// chart-boundary mesh vertices (:82-167), the crossings of every chart-boundary edge with its own
// chart's grid lines (:37-80), and on a seam the crossings of the opposite half-edge inserted by
// their parameter along the edge (:357-425); quadratic elements add edge midpoints shared by key
// (:525-603); a node's value is the mean over its appearances of the bilinear interpolation in
// the appearing chart (:700-760).  The reference triangulates each boundary-cell polygon with
// Triangle (a constrained Delaunay triangulation); here each convex piece (cell ∩ mesh triangle)
// that touches a seam is fanned from its centroid.  Both are conforming quadratic interpolants of
// the same node values along the seam; they differ only inside the piece.
#include "gdtp_boundary.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

namespace gdtp {

namespace {

// Degree-4 six-point rule on the unit triangle (weights sum to one).
const double kQP[6][2] = {{0.816847572980459, 0.0915762135097705}, {0.0915762135097705, 0.816847572980459},
                          {0.0915762135097705, 0.0915762135097705}, {0.108103018168070, 0.445948490915965},
                          {0.445948490915965, 0.108103018168070},  {0.445948490915965, 0.445948490915965}};
const double kQW[6] = {0.109951743655322, 0.109951743655322, 0.109951743655322,
                       0.223381589678011, 0.223381589678011, 0.223381589678011};

using Weights = NodeWeights;

// Seam pieces are fanned from their centroid; with GDTP_SEAM_FAN=vertex in the environment, from
// their first ring node instead.  Both are valid quadratic triangulations of the piece: tests
// compare the two to bound how much the solution depends on the choice (the reference's Triangle
// triangulation is a third one).
bool VertexFan() {
  static const bool v = [] {
    const char* e = std::getenv("GDTP_SEAM_FAN");
    return e && std::string(e) == "vertex";
  }();
  return v;
}

V2 ChartVertex(const Hierarchy& h, int c, int lv) { return h.atlas.charts[c].vertices[lv] - h.levels[0].charts[c].corner; }

// Parameters s in (0,1) where the segment a->b (chart-local coordinates of chart c) crosses a grid line.
void GridCrossings(const GridChart& g, V2 a, V2 b, double sFlip, std::vector<double>& out) {
  const double cs[2] = {g.cellW, g.cellH};
  for (int d = 0; d < 2; d++) {
    double lo = std::min(a[d], b[d]), hi = std::max(a[d], b[d]);
    if (hi - lo <= 0) continue;
    for (long m = (long)std::floor(lo / cs[d]); m <= (long)std::floor(hi / cs[d]); m++) {
      double s = (m * cs[d] - a[d]) / (b[d] - a[d]);
      if (s > 0 && s < 1) out.push_back(sFlip ? 1.0 - s : s);
    }
  }
}

// Bilinear interpolation weights of chart c at chart-local point q, on cell (pi,pj) when that cell
// contains q, else on the active cell containing q.
int Bilinear(const Hierarchy& h, int c, V2 q, int pi, int pj, double scale, Weights& w, std::string& err) {
  const Level& L0 = h.levels[0];
  const GridChart& g = L0.charts[c];
  const double tol = 1e-7;
  auto inside = [&](int ci, int cj) {
    if (ci < 0 || cj < 0 || ci >= g.width - 1 || cj >= g.height - 1) return false;
    double x = q.x / g.cellW - ci, y = q.y / g.cellH - cj;
    return x >= -tol && x <= 1 + tol && y >= -tol && y <= 1 + tol && g.localCell[(size_t)cj * (g.width - 1) + ci] >= 0;
  };
  int ci = -1, cj = -1;
  if (pi >= 0 && inside(pi, pj)) ci = pi, cj = pj;
  else {
    int fi = (int)std::floor(q.x / g.cellW), fj = (int)std::floor(q.y / g.cellH);
    for (int dj = 0; dj <= 1 && ci < 0; dj++)
      for (int di = 0; di <= 1 && ci < 0; di++)
        for (int sj = -1; sj <= 1 && ci < 0; sj += 2)
          for (int si = -1; si <= 1 && ci < 0; si += 2) {
            int ti = fi + (di ? si : 0), tj = fj + (dj ? sj : 0);
            if (inside(ti, tj)) ci = ti, cj = tj;
          }
  }
  if (ci < 0) { err = "seam node outside the active cells of its chart"; return 0; }
  const double x = std::min(1.0, std::max(0.0, q.x / g.cellW - ci)), y = std::min(1.0, std::max(0.0, q.y / g.cellH - cj));
  const double b[4] = {(1 - x) * (1 - y), x * (1 - y), x * y, (1 - x) * y};
  const auto& corners = g.cellIndices[g.localCell[(size_t)cj * (g.width - 1) + ci]];
  for (int k = 0; k < 4; k++) {
    if (std::fabs(b[k]) <= 1e-14) continue;
    int bd = L0.boundaryAndDeepIndex[corners[k]];
    if (bd <= 0) { err = "deep texel in the support of a boundary node"; return 0; }
    w.push_back({bd - 1, b[k] * scale});
  }
  return 1;
}

int VertexAverage(const SeamContext& ctx, int meshVertex, int c, int pi, int pj, Weights& w, std::string& err) {
  const auto& apps = ctx.vertexCharts[meshVertex];
  const double s = 1.0 / apps.size();
  for (const auto& a : apps)
    if (!Bilinear(*ctx.h, a.first, ChartVertex(*ctx.h, a.first, a.second), a.first == c ? pi : -1, pj, s, w, err)) return 0;
  return 1;
}

// Value of the seam point at parameter s along mesh half-edge mh, whose triangle lies in chart c.
int SeamValue(const SeamContext& ctx, int c, int mh, double s, int pi, int pj, Weights& w, std::string& err) {
  const Hierarchy& h = *ctx.h;
  const Atlas& A = h.atlas;
  const int mt = mh / 3, k = mh % 3;
  const AtlasChart& ac = A.charts[c];
  const int t = ctx.meshTriangleLocal[mt];
  const int la = ac.triangles[3 * t + k], lb = ac.triangles[3 * t + (k + 1) % 3];
  const double eps = 1e-12;
  if (s <= eps) return VertexAverage(ctx, ac.meshVertexIndices[la], c, pi, pj, w, err);
  if (s >= 1 - eps) return VertexAverage(ctx, ac.meshVertexIndices[lb], c, pi, pj, w, err);
  const V2 pa = ChartVertex(h, c, la), pb = ChartVertex(h, c, lb);
  if (!Bilinear(h, c, pa + (pb - pa) * s, pi, pj, 0.5, w, err)) return 0;
  // the opposite half-edge runs from this edge's end vertex to its start vertex
  const int mo = A.oppositeHalfEdge[mh], co = A.triangleChart[mo / 3], to = ctx.meshTriangleLocal[mo / 3], ko = mo % 3;
  const AtlasChart& ao = A.charts[co];
  const V2 qb = ChartVertex(h, co, ao.triangles[3 * to + ko]), qa = ChartVertex(h, co, ao.triangles[3 * to + (ko + 1) % 3]);
  return Bilinear(h, co, qa + (qb - qa) * s, -1, -1, 0.5, w, err);
}

// Quadratic-element mass and stiffness on triangle (p0,p1,p2); node order: corners 0,1,2, then the
// midpoints of (0,1), (1,2), (2,0).
void P2Element(const V2 p[3], const M2& Ginv, double gscale, double M[6][6], double S[6][6]) {
  const V2 e1 = p[1] - p[0], e2 = p[2] - p[0];
  const double det = cross(e1, e2);
  const double area = std::fabs(det) / 2.0 * gscale;
  // gradients of the barycentric coordinates (rows of the inverse Jacobian)
  const V2 g1(e2.y / det, -e2.x / det), g2(-e1.y / det, e1.x / det);
  const V2 gl[3] = {V2(-g1.x - g2.x, -g1.y - g2.y), g1, g2};
  for (int a = 0; a < 6; a++)
    for (int b = 0; b < 6; b++) M[a][b] = S[a][b] = 0;
  const int mi[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  for (int q = 0; q < 6; q++) {
    const double l[3] = {1 - kQP[q][0] - kQP[q][1], kQP[q][0], kQP[q][1]};
    double phi[6];
    V2 grad[6];
    for (int i = 0; i < 3; i++) phi[i] = l[i] * (2 * l[i] - 1), grad[i] = gl[i] * (4 * l[i] - 1);
    for (int m = 0; m < 3; m++) {
      const int i = mi[m][0], j = mi[m][1];
      phi[3 + m] = 4 * l[i] * l[j];
      grad[3 + m] = (gl[j] * l[i] + gl[i] * l[j]) * 4.0;
    }
    for (int a = 0; a < 6; a++) {
      const V2 ga = Ginv * grad[a];
      for (int b = 0; b < 6; b++) {
        M[a][b] += kQW[q] * phi[a] * phi[b] * area;
        S[a][b] += kQW[q] * dot(grad[b], ga) * area;
      }
    }
  }
}

}  // namespace

int BuildSeamContext(const Hierarchy& h, SeamContext& ctx, std::string& err) {
  (void)err;
  ctx.h = &h;
  const Atlas& A = h.atlas;
  const size_t nT = A.triangleChart.size();
  ctx.meshTriangleLocal.assign(nT, -1);
  int nV = 0;
  for (const auto& ac : A.charts) {
    for (size_t k = 0; k < ac.meshTriangleIndices.size(); k++) ctx.meshTriangleLocal[ac.meshTriangleIndices[k]] = (int)k;
    for (int v : ac.meshVertexIndices) nV = std::max(nV, v + 1);
  }
  ctx.seamHalfEdge.assign(3 * nT, 0);
  ctx.crossings.assign(3 * nT, {});
  ctx.vertexCharts.assign(nV, {});
  for (int c = 0; c < (int)A.charts.size(); c++) {
    const AtlasChart& ac = A.charts[c];
    for (int he : ac.boundaryHalfEdges) {
      const int t = he / 3, k = he % 3, mh = 3 * ac.meshTriangleIndices[t] + k;
      for (int e = 0; e < 2; e++) {
        const int lv = ac.triangles[3 * t + (k + e) % 3];
        auto& list = ctx.vertexCharts[ac.meshVertexIndices[lv]];
        if (std::find(list.begin(), list.end(), std::make_pair(c, lv)) == list.end()) list.push_back({c, lv});
      }
      if (A.oppositeHalfEdge[mh] >= 0) ctx.seamHalfEdge[mh] = 1;
    }
  }
  ctx.vertexChartCount.assign(nV, 0);
  for (int v = 0; v < nV; v++) ctx.vertexChartCount[v] = (int)ctx.vertexCharts[v].size();
  for (size_t mh = 0; mh < 3 * nT; mh++) {
    if (!ctx.seamHalfEdge[mh]) continue;
    const int c = A.triangleChart[mh / 3], t = ctx.meshTriangleLocal[mh / 3], k = mh % 3;
    const AtlasChart& ac = A.charts[c];
    const int mo = A.oppositeHalfEdge[mh], co = A.triangleChart[mo / 3], to = ctx.meshTriangleLocal[mo / 3], ko = mo % 3;
    const AtlasChart& ao = A.charts[co];
    auto& s = ctx.crossings[mh];
    GridCrossings(h.levels[0].charts[c], ChartVertex(h, c, ac.triangles[3 * t + k]), ChartVertex(h, c, ac.triangles[3 * t + (k + 1) % 3]),
                  0, s);
    GridCrossings(h.levels[0].charts[co], ChartVertex(h, co, ao.triangles[3 * to + ko]),
                  ChartVertex(h, co, ao.triangles[3 * to + (ko + 1) % 3]), 1, s);
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end(), [](double a, double b) { return b - a < 1e-12; }), s.end());
  }
  return 1;
}

int SeamPieceElements(const SeamContext& ctx, int c, int t, int i, int j, const std::vector<V2>& piece, std::vector<SeamElement>& els,
                      std::string& err) {
  const Hierarchy& h = *ctx.h;
  const AtlasChart& ac = h.atlas.charts[c];
  const GridChart& g = h.levels[0].charts[c];
  const int mt = ac.meshTriangleIndices[t];
  V2 P[3];
  int mv[3];
  bool seamEdge[3], seamVert[3], any = false;
  for (int k = 0; k < 3; k++) {
    P[k] = ChartVertex(h, c, ac.triangles[3 * t + k]);
    mv[k] = ac.meshVertexIndices[ac.triangles[3 * t + k]];
    seamEdge[k] = ctx.seamHalfEdge[3 * mt + k] != 0;
    seamVert[k] = ctx.vertexChartCount[mv[k]] > 1;
    any = any || seamEdge[k] || seamVert[k];
  }
  if (!any) return 0;
  const double tol = 1e-9 * std::min(g.cellW, g.cellH);
  // tags: -1 own chart; 0..2 triangle vertex k (a seam vertex); 3+k point on seam edge k at param s
  const int n = (int)piece.size();
  std::vector<int> tag(n, -1);
  std::vector<double> par(n, 0.0);
  bool touches = false;
  for (int v = 0; v < n; v++) {
    for (int k = 0; k < 3 && tag[v] < 0; k++) {
      const V2 d = piece[v] - P[k];
      if (seamVert[k] && std::sqrt(dot(d, d)) < tol) tag[v] = k;
    }
    for (int k = 0; k < 3 && tag[v] < 0; k++) {
      if (!seamEdge[k]) continue;
      const V2 a = P[k], e = P[(k + 1) % 3] - P[k];
      const double L2 = dot(e, e), s = dot(piece[v] - a, e) / L2;
      const double dist = std::fabs(cross(e, piece[v] - a)) / std::sqrt(L2);
      if (dist < tol && s > -1e-9 && s < 1 + 1e-9) tag[v] = 3 + k, par[v] = std::min(1.0, std::max(0.0, s));
    }
    touches = touches || tag[v] >= 0;
  }
  if (!touches) return 0;
  // parameter of a tagged point along seam edge k (NaN if it is not on that edge)
  auto onEdge = [&](int v, int k) -> double {
    if (tag[v] == 3 + k) return par[v];
    if (tag[v] == k) return 0.0;
    if (tag[v] == (k + 1) % 3) return 1.0;
    return NAN;
  };
  struct Node { V2 p; int kind; double s; };  // kind: -1 own, 0..2 vertex, 3+k seam edge k (s along it)
  struct RingEdge { int k; double s0, s1; };  // seam edge k (or -1) and the end parameters along it
  std::vector<Node> ring;
  std::vector<RingEdge> ringEdge;  // ring edge q joins ring[q] and ring[q+1]
  for (int v = 0; v < n; v++) {
    const int w = (v + 1) % n;
    ring.push_back({piece[v], tag[v], par[v]});
    int ek = -1;
    double sa = 0, sb = 0;
    for (int k = 0; k < 3 && ek < 0; k++) {
      if (!seamEdge[k]) continue;
      sa = onEdge(v, k), sb = onEdge(w, k);
      if (!std::isnan(sa) && !std::isnan(sb) && std::fabs(sb - sa) > 1e-14) ek = k;
    }
    if (ek < 0) { ringEdge.push_back({-1, 0, 0}); continue; }
    // insert the crossings of the seam edge that lie strictly between the end points
    const auto& cr = ctx.crossings[3 * mt + ek];
    const V2 a = P[ek], e = P[(ek + 1) % 3] - P[ek];
    std::vector<double> ins;
    const double lo = std::min(sa, sb), hi = std::max(sa, sb), gap = 1e-9;
    for (double s : cr)
      if (s > lo + gap && s < hi - gap) ins.push_back(s);
    if (sb < sa) std::reverse(ins.begin(), ins.end());
    double prev = sa;
    for (double s : ins) {
      ringEdge.push_back({ek, prev, s});
      ring.push_back({a + e * s, 3 + ek, s});
      prev = s;
    }
    ringEdge.push_back({ek, prev, sb});
  }
  // node weights
  auto weightsOf = [&](const Node& nd, Weights& w) -> int {
    if (nd.kind < 0) return Bilinear(h, c, nd.p, i, j, 1.0, w, err);
    if (nd.kind < 3) return VertexAverage(ctx, mv[nd.kind], c, i, j, w, err);
    return SeamValue(ctx, c, 3 * mt + (nd.kind - 3), nd.s, i, j, w, err);
  };
  const int m = (int)ring.size();
  std::vector<Weights> rw(m);
  for (int q = 0; q < m; q++)
    if (!weightsOf(ring[q], rw[q])) return -1;
  if (VertexFan()) {  // alternative triangulation: a fan from ring node 0 (quantifies the model's freedom)
    double area2 = 0.0;  // twice the piece's area; fan triangles of three collinear ring nodes are dropped
    for (int v = 0; v < n; v++) area2 += cross(piece[v], piece[(v + 1) % n]);
    for (int q = 1; q + 1 < m; q++) {
      SeamElement el;
      el.p[0] = ring[0].p, el.p[1] = ring[q].p, el.p[2] = ring[q + 1].p;
      if (std::fabs(cross(el.p[1] - el.p[0], el.p[2] - el.p[0])) < 1e-9 * std::fabs(area2)) continue;
      el.w[0] = rw[0], el.w[1] = rw[q], el.w[2] = rw[q + 1];
      // edge midpoints: ring edges carry their seam values, diagonals are inside the piece
      auto midWeights = [&](int a, int b, Weights& w) -> int {
        const int e = (b == a + 1) ? a : ((a == 0 && b == m - 1) ? m - 1 : -1);
        if (e >= 0 && ringEdge[e].k >= 0) {
          const Node mid{(ring[a].p + ring[b].p) * 0.5, 3 + ringEdge[e].k, 0.5 * (ringEdge[e].s0 + ringEdge[e].s1)};
          return weightsOf(mid, w);
        }
        return Bilinear(h, c, (ring[a].p + ring[b].p) * 0.5, i, j, 1.0, w, err);
      };
      if (!midWeights(0, q, el.w[3]) || !midWeights(q, q + 1, el.w[4]) || !midWeights(0, q + 1, el.w[5])) return -1;
      els.push_back(std::move(el));
    }
    return 1;
  }
  V2 cen(0, 0);
  for (const V2& p : piece) cen = cen + p;
  cen = cen / (double)n;
  Weights wc;
  if (!Bilinear(h, c, cen, i, j, 1.0, wc, err)) return -1;
  for (int q = 0; q < m; q++) {
    const int r = (q + 1) % m;
    SeamElement el;
    el.p[0] = cen, el.p[1] = ring[q].p, el.p[2] = ring[r].p;
    if (std::fabs(cross(el.p[1] - el.p[0], el.p[2] - el.p[0])) < 1e-24 * g.cellW * g.cellH) continue;
    el.w[0] = wc, el.w[1] = rw[q], el.w[2] = rw[r];
    if (!Bilinear(h, c, (cen + ring[q].p) * 0.5, i, j, 1.0, el.w[3], err)) return -1;
    if (ringEdge[q].k >= 0) {
      const Node mid{(ring[q].p + ring[r].p) * 0.5, 3 + ringEdge[q].k, 0.5 * (ringEdge[q].s0 + ringEdge[q].s1)};
      if (!weightsOf(mid, el.w[4])) return -1;
    } else if (!Bilinear(h, c, (ring[q].p + ring[r].p) * 0.5, i, j, 1.0, el.w[4], err)) return -1;
    if (!Bilinear(h, c, (ring[r].p + cen) * 0.5, i, j, 1.0, el.w[5], err)) return -1;
    els.push_back(std::move(el));
  }
  return 1;
}

int AssembleSeamPiece(const SeamContext& ctx, int c, int t, int i, int j, const std::vector<V2>& piece, const M2& metric,
                      std::vector<Triplet>& bbM, std::vector<Triplet>& bbS, std::string& err) {
  std::vector<SeamElement> els;
  const int rc = SeamPieceElements(ctx, c, t, i, j, piece, els, err);
  if (rc <= 0) return rc;
  // local dense accumulation over the boundary texels this piece touches
  std::vector<int> ids;
  auto local = [&](int b) {
    for (size_t q = 0; q < ids.size(); q++)
      if (ids[q] == b) return (int)q;
    ids.push_back(b);
    return (int)ids.size() - 1;
  };
  std::vector<double> LM, LS;
  auto grow = [&]() {
    const size_t nl = ids.size();
    if (LM.size() >= nl * nl) return;
    const size_t old = (size_t)std::sqrt((double)LM.size() + 0.5);
    std::vector<double> M2v(nl * nl, 0.0), S2v(nl * nl, 0.0);
    for (size_t r = 0; r < old; r++)
      for (size_t q = 0; q < old; q++) M2v[r * nl + q] = LM[r * old + q], S2v[r * nl + q] = LS[r * old + q];
    LM.swap(M2v), LS.swap(S2v);
  };
  const M2 Ginv = metric.inv();
  const double gscale = std::sqrt(metric.det());
  for (const SeamElement& el : els) {
    double Me[6][6], Se[6][6];
    P2Element(el.p, Ginv, gscale, Me, Se);
    std::vector<std::pair<int, double>> lw[6];
    for (int a = 0; a < 6; a++)
      for (const auto& x : el.w[a]) lw[a].push_back({local(x.first), x.second});
    grow();
    const size_t nl = ids.size();
    for (int a = 0; a < 6; a++)
      for (int b = 0; b < 6; b++)
        for (const auto& xa : lw[a])
          for (const auto& xb : lw[b]) {
            LM[xa.first * nl + xb.first] += Me[a][b] * xa.second * xb.second;
            LS[xa.first * nl + xb.first] += Se[a][b] * xa.second * xb.second;
          }
  }
  // emit the symmetric part (the two triangles of the local sums add in different orders)
  const size_t nl = ids.size();
  for (size_t r = 0; r < nl; r++)
    for (size_t q = 0; q < nl; q++) {
      const double m = 0.5 * (LM[r * nl + q] + LM[q * nl + r]), st = 0.5 * (LS[r * nl + q] + LS[q * nl + r]);
      if (m != 0.0) bbM.push_back({ids[r], ids[q], m});
      if (st != 0.0) bbS.push_back({ids[r], ids[q], st});
    }
  return 1;
}

int AssembleBoundaryCells(const Hierarchy& h, int c, const std::vector<std::array<double, 16>>& bcellM,
                          const std::vector<std::array<double, 16>>& bcellS, std::vector<Triplet>& bbM, std::vector<Triplet>& bbS,
                          std::string& err) {
  const Level& L0 = h.levels[0];
  const GridChart& g = L0.charts[c];
  for (int bc = 0; bc < g.numBoundaryCells; bc++) {
    const auto& gl = g.cellIndices[g.boundaryCellToLocal[bc]];
    int b[4];
    for (int k = 0; k < 4; k++) {
      b[k] = L0.boundaryAndDeepIndex[gl[k]] - 1;
      if (b[k] < 0) { err = "deep texel on a boundary cell"; return 0; }
    }
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < 4; l++) {
        bbM.push_back({b[k], b[l], bcellM[bc][4 * k + l]});
        bbS.push_back({b[k], b[l], bcellS[bc][4 * k + l]});
      }
  }
  return 1;
}

}  // namespace gdtp
