// Compressed-sparse-row matrices with sorted, duplicate-free rows (the row layout the
// reference's row-list SparseMatrix has after SparseMatrixParser, include/Src/SparseMatrixParser.h:60-116).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace gdtp {

struct Triplet {
  int r, c;
  double v;
};

struct Csr {
  int rows = 0, cols = 0;
  std::vector<int> rowptr{0}, col;
  std::vector<double> val;
  int64_t nnz() const { return (int64_t)col.size(); }

  // Sum duplicates, sort each row by column.
  static Csr FromTriplets(int rows, int cols, std::vector<Triplet>& t) {
    Csr m;
    m.rows = rows, m.cols = cols;
    std::stable_sort(t.begin(), t.end(), [](const Triplet& a, const Triplet& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
    m.rowptr.assign(rows + 1, 0);
    for (size_t k = 0; k < t.size();) {
      size_t e = k;
      double s = 0;
      while (e < t.size() && t[e].r == t[k].r && t[e].c == t[k].c) s += t[e].v, e++;
      m.col.push_back(t[k].c);
      m.val.push_back(s);
      m.rowptr[t[k].r + 1]++;
      k = e;
    }
    for (int i = 0; i < rows; i++) m.rowptr[i + 1] += m.rowptr[i];
    return m;
  }

  Csr Transpose() const {
    Csr t;
    t.rows = cols, t.cols = rows;
    t.rowptr.assign(cols + 1, 0);
    for (int c : col) t.rowptr[c + 1]++;
    for (int i = 0; i < cols; i++) t.rowptr[i + 1] += t.rowptr[i];
    t.col.resize(col.size());
    t.val.resize(val.size());
    std::vector<int> fill(t.rowptr.begin(), t.rowptr.end() - 1);
    for (int r = 0; r < rows; r++)
      for (int k = rowptr[r]; k < rowptr[r + 1]; k++) {
        int p = fill[col[k]]++;
        t.col[p] = r;
        t.val[p] = val[k];
      }
    return t;
  }

  // this * B (sparse-sparse, dense accumulator per row)
  Csr Multiply(const Csr& B) const {
    Csr C;
    C.rows = rows, C.cols = B.cols;
    C.rowptr.assign(rows + 1, 0);
    std::vector<double> acc(B.cols, 0.0);
    std::vector<int> mark(B.cols, -1), idx;
    for (int r = 0; r < rows; r++) {
      idx.clear();
      for (int k = rowptr[r]; k < rowptr[r + 1]; k++) {
        int j = col[k];
        double a = val[k];
        for (int q = B.rowptr[j]; q < B.rowptr[j + 1]; q++) {
          int c = B.col[q];
          if (mark[c] != r) mark[c] = r, acc[c] = 0, idx.push_back(c);
          acc[c] += a * B.val[q];
        }
      }
      std::sort(idx.begin(), idx.end());
      for (int c : idx) C.col.push_back(c), C.val.push_back(acc[c]);
      C.rowptr[r + 1] = (int)C.col.size();
    }
    return C;
  }

  // a*this + b*B (same shape)
  Csr Combine(double a, const Csr& B, double b) const {
    std::vector<Triplet> t;
    t.reserve(col.size() + B.col.size());
    for (int r = 0; r < rows; r++) {
      for (int k = rowptr[r]; k < rowptr[r + 1]; k++) t.push_back({r, col[k], a * val[k]});
      for (int k = B.rowptr[r]; k < B.rowptr[r + 1]; k++) t.push_back({r, B.col[k], b * B.val[k]});
    }
    return FromTriplets(rows, std::max(cols, B.cols), t);
  }

  // y (+)= this * x for `ncols` column-major vectors of length cols / rows
  void Apply(const double* x, double* y, int ncols, int ldx, int ldy, bool add, bool negate) const {
    for (int c = 0; c < ncols; c++)
      for (int r = 0; r < rows; r++) {
        double s = 0;
        for (int k = rowptr[r]; k < rowptr[r + 1]; k++) s += val[k] * x[(size_t)c * ldx + col[k]];
        if (negate) s = -s;
        if (add) y[(size_t)c * ldy + r] += s;
        else y[(size_t)c * ldy + r] = s;
      }
  }
};

}  // namespace gdtp
