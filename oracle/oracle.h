/* ORACLE — test infrastructure only.  Plain-C restatement of the reference V-cycle kernels
 * (include/Src/IterativeSolvers.inl), used by tests/ and by bench.py's cpu_baseline leg as the
 * checker; never linked into or called by the product library.
 *
 * Layout: vectors are column-major [K][N] doubles; deep stencils are the reference's AoS layout,
 * 10 doubles per deep texel (reciprocal form on smoothed levels: c_k = a_k/a_4, c_4 = 1/a_4,
 * c_9 = a_4; UpdateCoefficients.inl:374-397).  Segments are 5 ints
 * {currentStart, currentEnd, previousStart, nextStart, deepStart} (Hierarchy.h:81-104).
 */
#ifndef GDTP_ORACLE_H
#define GDTP_ORACLE_H

typedef struct {
  int n_tasks;
  const int* task_ptr;       /* n_tasks+1 -> blocks */
  const int* block_begin;    /* first deep row of block */
  const int* block_deep_end; /* end of deep rows == first padded row */
  const int* block_pad_end;  /* end of padded rows */
  const int* line_ptr;       /* row -> segments */
  const int* segments;       /* 5 ints per segment */
} oracle_tasks;

typedef struct {
  int rows;
  const int* rowptr;
  const int* col;
  const double* val;
} oracle_csr;

/* Deep part of Relaxation / RelaxationAndResidual: two-coloured band tasks
 * (IterativeSolvers.inl:497-521), residual into r when r != NULL. */
void oracle_band_relax(const oracle_tasks* t, const double* coef, double* x, const double* b, double* r, int N, int K);
/* vb[c][i] = b[c][bgi[i]] - (A_BD x)[i]  (IterativeSolvers.inl:403-419) */
void oracle_boundary_rhs(int B, const int* bgi, const oracle_csr* bd, const double* x, const double* b, double* vb, int N, int K);
/* r[c][bgi[i]] = b[c][bgi[i]] - (A_BD x)[i] - (A_BB x_B)[i]  (IterativeSolvers.inl:549-555) */
void oracle_boundary_residual(int B, const int* bgi, const oracle_csr* bd, const oracle_csr* bb, const double* x, const double* b,
                              double* r, int N, int K);
/* MultiplyByRestriction (IterativeSolvers.inl:1009-1071). */
void oracle_restrict(int Bc, const int* bgi_c, const oracle_csr* restr, int n_lines, const int* lines, const double* fine, int Nf,
                     double* coarse, int Nc, int K);
/* AddProlongation (IterativeSolvers.inl:1122-1168). */
void oracle_prolong_add(int n_lines, const int* lines, const double* coarse, int Nc, double* fine, int Nf, int K);
/* MultiplyBySystemMatrix_NoReciprocals (IterativeSolvers.inl:881-950); raw stencils. */
void oracle_apply(int n_raster, const int* raster, const double* coef_raw, int B, const int* bgi, const oracle_csr* bd,
                  const oracle_csr* bb, const double* x, double* y, int N, int K);

/* OpenMP threads of the parallel loops (iterations with disjoint outputs: any count gives the
 * sequential result bit for bit). */
void oracle_set_threads(int n);
int oracle_max_threads(void);

#endif
