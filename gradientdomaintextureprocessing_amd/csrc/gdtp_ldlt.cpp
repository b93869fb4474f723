#include "gdtp_ldlt.h"

#include <cmath>
#include <functional>

namespace gdtp {

// Recursive bisection at the middle breadth-first level from a pseudo-peripheral vertex; the
// separator is the part of that level adjacent to the next level, ordered last.
void NestedDissectionOrder(const Csr& A, std::vector<int>& perm) {
  const int n = A.rows;
  perm.clear();
  perm.reserve(n);
  std::vector<int> stamp(n, -1), level(n, -1), inSet(n, -1);
  int setId = 0, bfsId = 0;
  std::vector<int> queue;
  queue.reserve(n);
  auto bfs = [&](int src, int sid) {  // levels within set `sid`; returns vertices in BFS order
    queue.clear();
    int id = bfsId++;
    stamp[src] = id, level[src] = 0;
    queue.push_back(src);
    for (size_t h = 0; h < queue.size(); h++) {
      int v = queue[h];
      for (int k = A.rowptr[v]; k < A.rowptr[v + 1]; k++) {
        int u = A.col[k];
        if (u == v || inSet[u] != sid || stamp[u] == id) continue;
        stamp[u] = id, level[u] = level[v] + 1;
        queue.push_back(u);
      }
    }
  };
  // explicit work stack of vertex sets; a set is split into (A, B, S) and S is emitted after A and B.
  struct Item { std::vector<int> verts; bool emit; };
  std::vector<Item> stack;
  stack.push_back({std::vector<int>(n), false});
  for (int i = 0; i < n; i++) stack.back().verts[i] = i;
  const size_t kLeaf = 48;
  while (!stack.empty()) {
    Item it = std::move(stack.back());
    stack.pop_back();
    if (it.emit || it.verts.size() <= kLeaf) {
      perm.insert(perm.end(), it.verts.begin(), it.verts.end());
      continue;
    }
    int sid = setId++;
    for (int v : it.verts) inSet[v] = sid;
    bfs(it.verts[0], sid);
    if (queue.size() < it.verts.size()) {  // disconnected: split off this component
      std::vector<int> comp(queue.begin(), queue.end()), rest;
      int id = bfsId - 1;
      for (int v : it.verts)
        if (stamp[v] != id) rest.push_back(v);
      stack.push_back({std::move(rest), false});
      stack.push_back({std::move(comp), false});
      continue;
    }
    int far = queue.back();
    bfs(far, sid);
    far = queue.back();
    bfs(far, sid);
    int maxLevel = level[queue.back()];
    if (maxLevel < 2) {
      perm.insert(perm.end(), it.verts.begin(), it.verts.end());
      continue;
    }
    std::vector<int> cnt(maxLevel + 2, 0);
    for (int v : queue) cnt[level[v]]++;
    int half = (int)queue.size() / 2, acc = 0, mid = 1;
    for (int l = 0; l <= maxLevel; l++) {
      acc += cnt[l];
      if (acc >= half) { mid = std::max(1, std::min(l, maxLevel - 1)); break; }
    }
    std::vector<int> a, b, s;
    for (int v : queue) {
      int l = level[v];
      if (l < mid) a.push_back(v);
      else if (l > mid) b.push_back(v);
      else {
        bool touchesNext = false;
        for (int k = A.rowptr[v]; k < A.rowptr[v + 1]; k++) {
          int u = A.col[k];
          if (inSet[u] == sid && level[u] == mid + 1) { touchesNext = true; break; }
        }
        (touchesNext ? s : a).push_back(v);
      }
    }
    stack.push_back({std::move(s), true});
    stack.push_back({std::move(b), false});
    stack.push_back({std::move(a), false});
  }
}

int FactorLdlt(const Csr& A, Ldlt& F, std::string& err) {
  const int n = A.rows;
  F.n = n;
  NestedDissectionOrder(A, F.perm);
  F.iperm.assign(n, 0);
  for (int i = 0; i < n; i++) F.iperm[F.perm[i]] = i;
  // lower triangle of the permuted matrix by rows: C[k][i] (i <= k)
  std::vector<Triplet> t;
  t.reserve(A.col.size() / 2 + n);
  for (int r = 0; r < n; r++)
    for (int k = A.rowptr[r]; k < A.rowptr[r + 1]; k++) {
      int pr = F.iperm[r], pc = F.iperm[A.col[k]];
      if (pc <= pr) t.push_back({pr, pc, A.val[k]});
    }
  Csr C = Csr::FromTriplets(n, n, t);
  // symbolic: elimination tree and column counts
  std::vector<int> parent(n, -1), flag(n, -1), lnz(n, 0);
  for (int k = 0; k < n; k++) {
    flag[k] = k;
    for (int p = C.rowptr[k]; p < C.rowptr[k + 1]; p++)
      for (int i = C.col[p]; i < k && flag[i] != k; i = parent[i]) {
        if (parent[i] == -1) parent[i] = k;
        lnz[i]++;
        flag[i] = k;
      }
  }
  std::vector<int64_t> Lp(n + 1, 0);
  for (int k = 0; k < n; k++) Lp[k + 1] = Lp[k] + lnz[k];
  if (Lp[n] > (int64_t)2000000000) { err = "LDL^T factor too large"; return 0; }
  std::vector<int> Li(Lp[n]);
  std::vector<double> Lx(Lp[n]);
  std::vector<double> y(n, 0.0);
  std::vector<int> pattern(n), fill(n, 0);
  F.D.assign(n, 0.0);
  std::fill(flag.begin(), flag.end(), -1);
  // numeric: up-looking, row k of L by a sparse triangular solve over the etree reach
  for (int k = 0; k < n; k++) {
    int top = n;
    flag[k] = k;
    for (int p = C.rowptr[k]; p < C.rowptr[k + 1]; p++) {
      int i = C.col[p];
      y[i] += C.val[p];
      int len = 0;
      for (; flag[i] != k; i = parent[i]) pattern[len++] = i, flag[i] = k;
      while (len > 0) pattern[--top] = pattern[--len];
    }
    double d = y[k];
    y[k] = 0.0;
    for (; top < n; top++) {
      int i = pattern[top];
      double yi = y[i];
      y[i] = 0.0;
      int64_t p2 = Lp[i] + fill[i];
      for (int64_t p = Lp[i]; p < p2; p++) y[Li[p]] -= Lx[p] * yi;
      double lki = yi / F.D[i];
      d -= lki * yi;
      Li[p2] = k, Lx[p2] = lki;
      fill[i]++;
    }
    if (!(d > 0.0) || !std::isfinite(d)) { err = "matrix is not positive definite (LDL^T pivot " + std::to_string(d) + ")"; return 0; }
    F.D[k] = d;
  }
  // L by columns and by rows
  Csr& Lc = F.Lcols;
  Lc.rows = n, Lc.cols = n;
  Lc.rowptr.assign(n + 1, 0);
  for (int k = 0; k < n; k++) Lc.rowptr[k + 1] = (int)Lp[k + 1];
  Lc.col = Li, Lc.val = Lx;
  F.Lrows = Lc.Transpose();
  // dependency levels
  std::vector<int> lev(n, 0);
  int maxL = 0;
  for (int k = 0; k < n; k++) {
    int l = 0;
    for (int p = F.Lrows.rowptr[k]; p < F.Lrows.rowptr[k + 1]; p++) l = std::max(l, lev[F.Lrows.col[p]] + 1);
    lev[k] = l, maxL = std::max(maxL, l);
  }
  auto bucket = [&](std::vector<int>& ptr, std::vector<int>& rows, int nl) {
    ptr.assign(nl + 2, 0);
    for (int k = 0; k < n; k++) ptr[lev[k] + 1]++;
    for (int l = 0; l <= nl; l++) ptr[l + 1] += ptr[l];
    rows.assign(n, 0);
    std::vector<int> f(ptr.begin(), ptr.end() - 1);
    for (int k = 0; k < n; k++) rows[f[lev[k]]++] = k;
    ptr.resize(nl + 2);
  };
  bucket(F.fwdLevelPtr, F.fwdRows, maxL);
  maxL = 0;
  for (int i = n - 1; i >= 0; i--) {
    int l = 0;
    for (int p = Lc.rowptr[i]; p < Lc.rowptr[i + 1]; p++) l = std::max(l, lev[Lc.col[p]] + 1);
    lev[i] = l, maxL = std::max(maxL, l);
  }
  bucket(F.bwdLevelPtr, F.bwdRows, maxL);
  return 1;
}

void SolveLdlt(const Ldlt& F, const double* b, double* x, int ncols, int ld) {
  const int n = F.n;
  std::vector<double> y(n);
  for (int c = 0; c < ncols; c++) {
    for (int k = 0; k < n; k++) {
      double s = b[(size_t)c * ld + F.perm[k]];
      for (int p = F.Lrows.rowptr[k]; p < F.Lrows.rowptr[k + 1]; p++) s -= F.Lrows.val[p] * y[F.Lrows.col[p]];
      y[k] = s;
    }
    for (int k = 0; k < n; k++) y[k] /= F.D[k];
    for (int i = n - 1; i >= 0; i--) {
      double s = y[i];
      for (int p = F.Lcols.rowptr[i]; p < F.Lcols.rowptr[i + 1]; p++) s -= F.Lcols.val[p] * y[F.Lcols.col[p]];
      y[i] = s;
    }
    for (int k = 0; k < n; k++) x[(size_t)c * ld + F.perm[k]] = y[k];
  }
}

}  // namespace gdtp
