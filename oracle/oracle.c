/* ORACLE — test infrastructure only (see oracle.h).  Each function states the reference lines
 * it restates; loops run in the reference's sequential order, or in parallel (OpenMP) over
 * iterations that write disjoint outputs, which gives the sequential result bit for bit. */
#include "oracle.h"

#include <omp.h>
#include <stddef.h>

/* Threads of the parallel loops (tests compare one thread with many: the results are identical). */
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }

/* UpdateRow (IterativeSolvers.inl:438-465): left-to-right Gauss-Seidel over the row's segments. */
static void update_row(const oracle_tasks* t, int row, const double* coef, double* x, const double* b, int N, int K) {
  for (int s = t->line_ptr[row]; s < t->line_ptr[row + 1]; s++) {
    const int* S = t->segments + 5 * s;
    int len = S[1] - S[0] + 1;
    for (int k = 0; k < K; k++) {
      double* xc = x + (size_t)k * N + S[0];
      const double* xp = x + (size_t)k * N + S[2];
      const double* xn = x + (size_t)k * N + S[3];
      const double* rh = b + (size_t)k * N + S[0];
      const double* c = coef + 10 * (size_t)S[4];
      for (int i = 0; i < len; i++, c += 10)
        xc[i] = rh[i] * c[4] - (xp[i - 1] * c[0] + xp[i] * c[1] + xp[i + 1] * c[2] + xc[i - 1] * c[3] + xc[i + 1] * c[5] +
                                xn[i - 1] * c[6] + xn[i] * c[7] + xn[i + 1] * c[8]);
    }
  }
}

/* UpdateResidual (IterativeSolvers.inl:467-495). */
static void update_residual(const oracle_tasks* t, int row, const double* coef, const double* x, const double* b, double* r, int N,
                            int K) {
  for (int s = t->line_ptr[row]; s < t->line_ptr[row + 1]; s++) {
    const int* S = t->segments + 5 * s;
    int len = S[1] - S[0] + 1;
    for (int k = 0; k < K; k++) {
      const double* xc = x + (size_t)k * N + S[0];
      const double* xp = x + (size_t)k * N + S[2];
      const double* xn = x + (size_t)k * N + S[3];
      const double* rh = b + (size_t)k * N + S[0];
      double* rr = r + (size_t)k * N + S[0];
      const double* c = coef + 10 * (size_t)S[4];
      for (int i = 0; i < len; i++, c += 10)
        rr[i] = rh[i] - (xp[i - 1] * c[0] + xp[i] * c[1] + xp[i + 1] * c[2] + xc[i - 1] * c[3] + xc[i] + xc[i + 1] * c[5] +
                         xn[i - 1] * c[6] + xn[i] * c[7] + xn[i + 1] * c[8]) *
                            c[9];
    }
  }
}

/* UpdateBlock with 3 lagged passes (IterativeSolvers.inl:497-509). */
static void update_block(const oracle_tasks* t, int blk, const double* coef, double* x, const double* b, double* r, int N, int K) {
  const int gs = 3;
  int r0 = t->block_begin[blk], p0 = t->block_deep_end[blk];
  int nDeep = p0 - r0, nPad = t->block_pad_end[blk] - p0;
  for (int j = 0; j < nDeep + gs; j++) {
    for (int g = 0; g < gs; g++)
      if (j - g >= 0 && j - g < nDeep) update_row(t, r0 + j - g, coef, x, b, N, K);
    if (r && j - gs >= 0 && j - gs < nPad) update_residual(t, p0 + j - gs, coef, x, b, r, N, K);
  }
  if (r)
    for (int j = nDeep; j < nPad; j++) update_residual(t, p0 + j, coef, x, b, r, N, K);
}

void oracle_band_relax(const oracle_tasks* t, const double* coef, double* x, const double* b, double* r, int N, int K) {
  /* even tasks, then odd tasks (IterativeSolvers.inl:511-521); the tasks of one colour touch
   * disjoint rows, so they run in parallel as in the reference, with the sequential result */
  for (int color = 0; color < 2; color++) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int task = color; task < t->n_tasks; task += 2)
      for (int blk = t->task_ptr[task]; blk < t->task_ptr[task + 1]; blk++) update_block(t, blk, coef, x, b, r, N, K);
  }
}

static double row_dot(const oracle_csr* m, int i, const double* x) {
  double s = 0;
  for (int k = m->rowptr[i]; k < m->rowptr[i + 1]; k++) s += x[m->col[k]] * m->val[k];
  return s;
}

void oracle_boundary_rhs(int B, const int* bgi, const oracle_csr* bd, const double* x, const double* b, double* vb, int N, int K) {
  for (int k = 0; k < K; k++)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < B; i++) vb[(size_t)k * B + i] = b[(size_t)k * N + bgi[i]] - row_dot(bd, i, x + (size_t)k * N);
}

void oracle_boundary_residual(int B, const int* bgi, const oracle_csr* bd, const oracle_csr* bb, const double* x, const double* b,
                              double* r, int N, int K) {
  for (int k = 0; k < K; k++) {
    const double* X = x + (size_t)k * N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < B; i++) {
      double v = b[(size_t)k * N + bgi[i]] - row_dot(bd, i, X);
      double s = 0;
      for (int q = bb->rowptr[i]; q < bb->rowptr[i + 1]; q++) s += X[bgi[bb->col[q]]] * bb->val[q];
      r[(size_t)k * N + bgi[i]] = v - s;
    }
  }
}

void oracle_restrict(int Bc, const int* bgi_c, const oracle_csr* restr, int n_lines, const int* lines, const double* fine, int Nf,
                     double* coarse, int Nc, int K) {
  for (int k = 0; k < K; k++) {
    const double* in = fine + (size_t)k * Nf;
    double* out = coarse + (size_t)k * Nc;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Bc; i++) out[bgi_c[i]] = row_dot(restr, i, in);
#pragma omp parallel for schedule(static)
    for (int l = 0; l < n_lines; l++) {
      const int* L = lines + 5 * l;
      const double *cu = in + L[0], *pv = in + L[2], *nx = in + L[3];
      double* o = out + L[4];
      int len = (L[1] - L[0]) / 2 + 1;
      for (int i = 0, j = 0; i < len; i++, j += 2)
        o[i] = cu[j] + (pv[j] + cu[j - 1] + cu[j + 1] + nx[j]) * 0.5 + (pv[j - 1] + pv[j + 1] + nx[j - 1] + nx[j + 1]) * 0.25;
    }
  }
}

void oracle_prolong_add(int n_lines, const int* lines, const double* coarse, int Nc, double* fine, int Nf, int K) {
  for (int k = 0; k < K; k++) {
    const double* in = coarse + (size_t)k * Nc;
    double* out = fine + (size_t)k * Nf;
#pragma omp parallel for schedule(static)
    for (int l = 0; l < n_lines; l++) {
      const int* L = lines + 6 * l;
      int off = L[5] ? 0 : 1;
      for (int i = 0; i < L[1]; i++) {
        int p = (i + off) / 2, q = (i + off + 1) / 2;
        if (L[2] != -1) out[L[0] + i] += (in[L[2] + p] + in[L[2] + q]) * 0.5;
        else out[L[0] + i] += (in[L[3] + p] + in[L[3] + q] + in[L[4] + p] + in[L[4] + q]) * 0.25;
      }
    }
  }
}

void oracle_apply(int n_raster, const int* raster, const double* coef_raw, int B, const int* bgi, const oracle_csr* bd,
                  const oracle_csr* bb, const double* x, double* y, int N, int K) {
  for (int k = 0; k < K; k++) {
    const double* X = x + (size_t)k * N;
    double* Y = y + (size_t)k * N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < B; i++) {
      double s = 0;
      for (int q = bb->rowptr[i]; q < bb->rowptr[i + 1]; q++) s += X[bgi[bb->col[q]]] * bb->val[q];
      Y[bgi[i]] = s + row_dot(bd, i, X);
    }
#pragma omp parallel for schedule(static)
    for (int l = 0; l < n_raster; l++) {
      const int* L = raster + 5 * l;
      const double *cu = X + L[0], *pv = X + L[2], *nx = X + L[3];
      const double* c = coef_raw + 10 * (size_t)L[4];
      for (int i = 0; i <= L[1] - L[0]; i++, c += 10)
        Y[L[0] + i] = pv[i - 1] * c[0] + pv[i] * c[1] + pv[i + 1] * c[2] + cu[i - 1] * c[3] + cu[i] * c[4] + cu[i + 1] * c[5] +
                      nx[i - 1] * c[6] + nx[i] * c[7] + nx[i + 1] * c[8];
    }
  }
}
