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
    // boundary-deep rows from the coarse deep stencils by symmetry (UpdateCoefficients.inl:89-101)
    std::vector<Triplet> t;
    t.reserve(C.boundaryDeepIndices.size());
    for (const auto& x : C.boundaryDeepIndices) t.push_back({x.boundary, x.deepGlobal, co.deepRaw[10 * (size_t)x.deep + x.offset]});
    co.bd = Csr::FromTriplets(C.numBoundary, C.numTexels, t);
    // boundary-boundary: R A_bb P + couplings through fine deep texels (UpdateCoefficients.inl:104-135)
    Csr partial = h.levels[l - 1].boundaryFineCoarseRestriction.Multiply(fo.bb.Multiply(C.boundaryCoarseFineProlongation));
    t.clear();
    t.reserve(C.boundaryBoundaryIndices.size());
    for (const auto& x : C.boundaryBoundaryIndices)
      t.push_back({x.principal, x.secondary, fo.deepRaw[10 * (size_t)x.fineDeep + x.offset] * x.weight});
    Csr T = Csr::FromTriplets(C.numBoundary, C.numBoundary, t);
    co.bb = T.Combine(1.0, partial, 1.0);
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
