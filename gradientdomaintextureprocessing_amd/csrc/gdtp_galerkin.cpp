#include "gdtp_galerkin.h"

#include <omp.h>

namespace gdtp {

// One-dimensional bilinear prolongation weight of a fine offset d from a coarse node: 1 at 0,
// 1/2 at +-1, 0 beyond.
static inline double Hat(int d) { return d == 0 ? 1.0 : ((d == 1 || d == -1) ? 0.5 : 0.0); }

void RestrictDeepCoefficients(const std::vector<double>& fine, std::vector<double>& coarse, const std::vector<DeepLine>& lines) {
  // Coarse stencil entry toward coarse neighbour E:  sum over fine texels f within one texel of the
  // coarse centre (prolongation weight w_f) and their stencil entries e of  w_f * A_f[e] * P(f+e, E).
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)lines.size(); r++) {
    const DeepLine& dl = lines[r];
    const int len = dl.coarseEnd - dl.coarseStart + 1;
    for (int i = 0; i < len; i++) {
      double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int fy = -1; fy <= 1; fy++) {
        int rowDeep = fy < 0 ? dl.finePrev : (fy > 0 ? dl.fineNext : dl.fineCurrent);
        for (int fx = -1; fx <= 1; fx++) {
          const double* a = &fine[10 * ((size_t)rowDeep + 2 * i + fx)];
          double wf = Hat(fx) * Hat(fy);
          for (int ey = -1; ey <= 1; ey++)
            for (int ex = -1; ex <= 1; ex++) {
              double v = wf * a[3 * (ey + 1) + (ex + 1)];
              int gx = fx + ex, gy = fy + ey;
              for (int Ey = -1; Ey <= 1; Ey++) {
                double py = Hat(gy - 2 * Ey);
                if (py == 0) continue;
                for (int Ex = -1; Ex <= 1; Ex++) {
                  double px = Hat(gx - 2 * Ex);
                  if (px != 0) acc[3 * (Ey + 1) + (Ex + 1)] += v * px * py;
                }
              }
            }
        }
      }
      double* c = &coarse[10 * ((size_t)dl.coarseStart + i)];
      for (int k = 0; k < 9; k++) c[k] = acc[k];
      c[9] = 0;
    }
  }
}

// Coarse parents of fine texel f with their bilinear weights (row f of the prolongation): per
// axis, the coarse texel at floor(u / 2) and, for odd u, also the next one, each with weight 1/2,
// u being f's chart-grid offset from the chart centre (the level grids are aligned at chart
// centres, ProlongationAndRestriction.inl:99-172).  Returns the number of parents.
static int ProlongRow(const Level& F, const Level& C, int f, int (&col)[4], double (&w)[4]) {
  const NodeInfo& ni = F.nodeInfo[f];
  const GridChart &fc = F.charts[ni.chart], &cc = C.charts[ni.chart];
  const int ui = ni.ci - fc.centerOffset[0], uj = ni.cj - fc.centerOffset[1];
  auto floorHalf = [](int u) { return u >= 0 ? u / 2 : -((1 - u) / 2); };
  const int ci = floorHalf(ui) + cc.centerOffset[0], cj = floorHalf(uj) + cc.centerOffset[1];
  const double wt = ((ui & 1) ? 0.5 : 1.0) * ((uj & 1) ? 0.5 : 1.0);
  int n = 0;
  for (int dj = 0; dj <= (uj & 1); dj++)
    for (int di = 0; di <= (ui & 1); di++) col[n] = cc.texel(ci + di, cj + dj), w[n] = wt, n++;
  return n;
}

// Calls visit(g, a) for the entries a = A[f][g] of row f of a level's operator: the nine-point
// stencil of a deep texel, or the A_BD and A_BB rows of a boundary texel.
template <class Visit>
static void OperatorRow(const Level& L, const LevelOperator& o, int f, Visit&& visit) {
  const int bd = L.boundaryAndDeepIndex[f];
  if (bd < 0) {
    const int d = -bd - 1;
    const NodeInfo& ni = L.nodeInfo[f];
    const GridChart& c = L.charts[ni.chart];
    for (int k = 0; k < 9; k++) visit(c.texel(ni.ci + k % 3 - 1, ni.cj + k / 3 - 1), o.deepRaw[10 * (size_t)d + k]);
  } else {
    const int b = bd - 1;
    for (int k = o.bd.rowptr[b]; k < o.bd.rowptr[b + 1]; k++) visit(o.bd.col[k], o.bd.val[k]);
    for (int k = o.bb.rowptr[b]; k < o.bb.rowptr[b + 1]; k++) visit(L.boundaryGlobalIndex[o.bb.col[k]], o.bb.val[k]);
  }
}

// Boundary rows of the coarse Galerkin operator P^T A P (UpdateCoefficients.inl:89-135 forms the
// same rows from index lists): row i of the coarse boundary is (R_B A) P with R_B = rows of P^T at
// the coarse boundary texels (the fine level's boundaryRestriction), evaluated row by row over
// the few fine texels each row touches.  Coarse columns that are deep go to A_BD (global
// columns), boundary ones to A_BB (boundary columns).
static void CoarseBoundaryRows(const Level& F, const LevelOperator& fo, const Level& C, Csr& bd, Csr& bb) {
  const int B = C.numBoundary;
  std::vector<std::vector<std::pair<int, double>>> rows(B);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < B; i++) {
    std::vector<std::pair<int, double>>& acc = rows[i];
    auto add = [&](int c, double v) {
      for (auto& e : acc)
        if (e.first == c) {
          e.second += v;
          return;
        }
      acc.push_back({c, v});
    };
    const Csr& R = F.boundaryRestriction;
    for (int k = R.rowptr[i]; k < R.rowptr[i + 1]; k++) {
      const int f = R.col[k];
      const double rw = R.val[k];
      OperatorRow(F, fo, f, [&](int g, double a) {
        int pc[4];
        double pw[4];
        const int n = ProlongRow(F, C, g, pc, pw);
        for (int q = 0; q < n; q++) add(pc[q], rw * a * pw[q]);
      });
    }
    std::sort(acc.begin(), acc.end());
  }
  std::vector<Triplet> td, tb;
  for (int i = 0; i < B; i++)
    for (const auto& e : rows[i]) {
      const int x = C.boundaryAndDeepIndex[e.first];
      if (x < 0) td.push_back({i, e.first, e.second});
      else tb.push_back({i, x - 1, e.second});
    }
  bd = Csr::FromTriplets(B, C.numTexels, td);
  bb = Csr::FromTriplets(B, B, tb);
}

Csr FullLevelMatrix(const Level& L, const std::vector<double>& deepRaw, const Csr& bb, const Csr& bd) {
  std::vector<Triplet> t;
  t.reserve((size_t)L.numDeep * 9 + bb.col.size() + bd.col.size() + L.numTexels);
  for (const Line5& rl : L.rasterLines) {
    const int len = rl.end - rl.start + 1;
    for (int i = 0; i < len; i++) {
      const double* a = &deepRaw[10 * ((size_t)rl.coeff + i)];
      const int row = rl.start + i, base[3] = {rl.prev + i, rl.start + i, rl.next + i};
      for (int dy = 0; dy < 3; dy++)
        for (int dx = 0; dx < 3; dx++) t.push_back({row, base[dy] + dx - 1, a[3 * dy + dx]});
    }
  }
  for (int b = 0; b < L.numBoundary; b++) {
    int row = L.boundaryGlobalIndex[b];
    for (int k = bd.rowptr[b]; k < bd.rowptr[b + 1]; k++) t.push_back({row, bd.col[k], bd.val[k]});
    for (int k = bb.rowptr[b]; k < bb.rowptr[b + 1]; k++) t.push_back({row, L.boundaryGlobalIndex[bb.col[k]], bb.val[k]});
  }
  for (int i = 0; i < L.numTexels; i++) t.push_back({i, i, 0.0});  // ensure a diagonal entry
  return Csr::FromTriplets(L.numTexels, L.numTexels, t);
}

int BuildOperatorHierarchy(const Hierarchy& h, const FineOperator& op, double screen, double stiff, bool factor,
                           OperatorHierarchy& oh, std::string& err) {
  const int nl = (int)h.levels.size();
  oh.levels.assign(nl, LevelOperator());
  LevelOperator& f = oh.levels[0];
  f.deepRaw.resize(op.deepMass.size());
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)f.deepRaw.size(); i++) f.deepRaw[i] = op.deepMass[i] * screen + op.deepStiffness[i] * stiff;
  f.bb = op.bbMass.Combine(screen, op.bbStiffness, stiff);
  f.bd = op.bdMass.Combine(screen, op.bdStiffness, stiff);
  for (int l = 1; l < nl; l++) {
    const Level& C = h.levels[l];
    const LevelOperator& fo = oh.levels[l - 1];
    LevelOperator& co = oh.levels[l];
    // coarse deep lines need the 3x3 fine neighbourhood of every coarse deep texel to be deep
    for (const DeepLine& dl : C.deepLines) {
      int len = dl.coarseEnd - dl.coarseStart + 1;
      int nFineDeep = (int)fo.deepRaw.size() / 10;
      for (int base : {dl.finePrev, dl.fineCurrent, dl.fineNext})
        if (base - 1 < 0 || base + 2 * (len - 1) + 1 >= nFineDeep) { err = "coarse deep line leaves the fine deep range"; return 0; }
    }
    co.deepRaw.assign(10 * (size_t)C.numDeep, 0.0);
    RestrictDeepCoefficients(fo.deepRaw, co.deepRaw, C.deepLines);
    CoarseBoundaryRows(h.levels[l - 1], fo, C, co.bd, co.bb);
  }
  oh.coarseMatrix = FullLevelMatrix(h.levels[nl - 1], oh.levels[nl - 1].deepRaw, oh.levels[nl - 1].bb, oh.levels[nl - 1].bd);
  // reciprocal form of the smoothed levels (UpdateCoefficients.inl:374-397)
  for (int l = 0; l < nl; l++) {
    LevelOperator& o = oh.levels[l];
    o.deepRecip = o.deepRaw;
    if (l == nl - 1) continue;
    const long nd = (long)o.deepRaw.size() / 10;
#pragma omp parallel for schedule(static)
    for (long d = 0; d < nd; d++) {
      double* c = &o.deepRecip[10 * d];
      double a4 = c[4], r = 1.0 / a4;
      for (int k = 0; k < 9; k++) c[k] *= r;
      c[4] = r;
      c[9] = a4;
    }
  }
  if (factor) {
    oh.boundaryFactor.assign(nl > 1 ? nl - 1 : 0, Ldlt());
    for (int l = 0; l + 1 < nl; l++)
      if (!FactorLdlt(oh.levels[l].bb, oh.boundaryFactor[l], err)) { err = "boundary level " + std::to_string(l) + ": " + err; return 0; }
    if (!FactorLdlt(oh.coarseMatrix, oh.coarseFactor, err)) { err = "coarse level: " + err; return 0; }
  }
  return 1;
}

}  // namespace gdtp
