// HIP kernels of the texel multigrid (CDNA4 / gfx950) and their host orchestration.
//
// Vectors are column-major: column c of a level-l vector starts at c*ld.  Deep stencils are
// stored structure-of-arrays, [10][D], so every stencil slot is a coalesced stream.
#include <stdexcept>

#include "gdtp_device.h"

#include <mutex>

namespace gdtp {

static constexpr int kWave = 64;
static constexpr int kDenseTopMax = 4096;  // largest LDL^T top solved through a dense inverse

// ------------------------------------------------------------------ band Gauss-Seidel
// The reference smoother (IterativeSolvers.inl:398-558 with residual, :174-395 without):
// per band task, blocks left to right; inside a block, three lagged passes over the block's deep
// rows (at step j rows j, j-1, j-2), each row a left-to-right Gauss-Seidel sweep of its deep
// segments, and (pre-smoother) the true residual of padded row j-3.
// The in-row dependence x_i = beta_i - c3_i x_{i-1} is a first-order linear recurrence, evaluated
// with a wavefront-wide scan (two texels per lane, 128-texel rows).

// Where one lane's texels of a row are: up to three elements (a pass takes elements 2 lane and
// 2 lane + 1, for the scan; a residual row lane, lane + 64 and lane + 128), each with its global
// index g, the rows above and below (p, n), its deep index d and whether it starts a segment.  The
// host expands every band line into a table of such elements (BandTasks::elems), so that a wave
// finds its elements with one load that waits on nothing, issued a step ahead (LoadRow); searching
// the line's segments instead took three or four dependent memory round trips per step.
struct RowLoc {
  int g[3], p[3], n[3], d[3];
  bool first[3];
};

// One left-to-right Gauss-Seidel sweep of a row (UpdateRow, IterativeSolvers.inl:448-470): the
// in-row dependence x_i = beta_i - c3_i x_{i-1} as a wavefront scan over the lanes' element pairs,
// columns in chunks of kKc (registers: the chunk's values and scan state).
template <int kKc>
__device__ __forceinline__ void GsRow(const RowLoc& L, const double* __restrict__ coef, int D, double* x, const double* __restrict__ b,
                                      int K, int ld) {
  const int lane = threadIdx.x % kWave;
  double a[2] = {0, 0}, c[2][9];
  for (int q = 0; q < 2; q++) {
    if (L.g[q] >= 0) {
      for (int k = 0; k < 9; k++) c[q][k] = coef[(size_t)k * D + L.d[q]];
      a[q] = L.first[q] ? 0.0 : -c[q][3];
    }
  }
  for (int c0 = 0; c0 < K; c0 += kKc) {
    const int kc = min(kKc, K - c0);
    double beta[2][kKc];
    for (int q = 0; q < 2; q++)
      for (int cc = 0; cc < kKc; cc++) {
        beta[q][cc] = 0.0;
        if (L.g[q] < 0 || cc >= kc) continue;
        const double* X = x + (size_t)(c0 + cc) * ld;
        const double* B = b + (size_t)(c0 + cc) * ld;
        const int g = L.g[q], p = L.p[q], n = L.n[q];
        double s = X[p - 1] * c[q][0] + X[p] * c[q][1] + X[p + 1] * c[q][2] + X[g + 1] * c[q][5] + X[n - 1] * c[q][6] + X[n] * c[q][7] +
                   X[n + 1] * c[q][8];
        double v = B[g] * c[q][4] - s;
        if (L.first[q]) v -= X[g - 1] * c[q][3];
        beta[q][cc] = v;
      }
    // lane-local composition of the two elements, then an inclusive wavefront scan
    double A = a[1] * a[0];
    double Bs[kKc];
    for (int cc = 0; cc < kKc; cc++) Bs[cc] = a[1] * beta[0][cc] + beta[1][cc];
    for (int o = 1; o < kWave; o <<= 1) {
      double Au = __shfl_up(A, o, kWave);
      double Bu[kKc];
      for (int cc = 0; cc < kKc; cc++) Bu[cc] = __shfl_up(Bs[cc], o, kWave);
      if (lane >= o) {
        for (int cc = 0; cc < kKc; cc++) Bs[cc] = A * Bu[cc] + Bs[cc];
        A = A * Au;
      }
    }
    for (int cc = 0; cc < kKc; cc++) {
      double prev = __shfl_up(Bs[cc], 1, kWave);
      if (lane == 0) prev = 0.0;
      double x0 = a[0] * prev + beta[0][cc];
      double x1 = a[1] * x0 + beta[1][cc];
      if (cc < kc) {
        double* X = x + (size_t)(c0 + cc) * ld;
        if (L.g[0] >= 0) X[L.g[0]] = x0;
        if (L.g[1] >= 0) X[L.g[1]] = x1;
      }
    }
  }
}

// The true residual of a padded row (UpdateResidual, IterativeSolvers.inl:473-495).
__device__ __forceinline__ void ResRow(const RowLoc& L, const double* __restrict__ coef, int D, const double* x,
                                       const double* __restrict__ b, double* r, int K, int ld) {
  for (int q = 0; q < 3; q++) {
    if (L.g[q] < 0) continue;
    const int g = L.g[q], p = L.p[q], n = L.n[q];
    double c[10];
    for (int k = 0; k < 10; k++) c[k] = coef[(size_t)k * D + L.d[q]];
    for (int cc = 0; cc < K; cc++) {
      const double* X = x + (size_t)cc * ld;
      double s = X[p - 1] * c[0] + X[p] * c[1] + X[p + 1] * c[2] + X[g - 1] * c[3] + X[g] + X[g + 1] * c[5] + X[n - 1] * c[6] +
                 X[n] * c[7] + X[n + 1] * c[8];
      r[(size_t)cc * ld + g] = b[(size_t)cc * ld + g] - s * c[9];
    }
  }
}

// One workgroup per band task.  Within a block the three lagged passes and the residual rows run as
// a pipeline, one row per wave per step: at local step t wave 0 relaxes row t (first pass), wave 1
// row t-2 (second pass), wave 2 row t-4 (third pass) and wave 3 computes the residual of padded row
// t-6.  Each operation sees exactly the values it sees in the reference's order (step j: rows j,
// j-1, j-2, then padded row j-3, IterativeSolvers.inl:497-509): a pass reads its row's neighbours
// above after their own pass and below before it, which the two-step lag preserves, and a padded
// row's residual reads deep rows up to one past its own index, final after their third pass (done
// by step pr+5).  Rows of one step are disjoint.
// With G = 1 the blocks of a task (left to right, overlapping by two columns,
// IterativeSolvers.inl:511-515) run one after another.  With G = 3 they also overlap in time: three
// groups of four waves take blocks in turn, block b running its local steps at task steps
// blockStart[b] + t; the host schedule (BandSchedule) starts a block only once no chart row it
// touches (reads or writes, rows and residuals, one row either side) is still touched by the block
// before it, so every value is read and written in the reference's order.  G = 3 pays where a
// launch has fewer tasks than CUs and several blocks per task; with more tasks, the 12-wave
// workgroups (about 160 registers a lane) fit one per CU and the launch takes several rounds.
struct BandBlock {  // one block's rows as the band kernel walks them
  int r0 = 0, p0 = 0, nDeep = 0, nPad = 0, steps = 0;
};
__device__ __forceinline__ BandBlock LoadBlock(const int* __restrict__ blockBegin, const int* __restrict__ blockDeepEnd,
                                               const int* __restrict__ blockPadEnd, int blk, int withResidual) {
  BandBlock B;
  B.r0 = blockBegin[blk], B.p0 = blockDeepEnd[blk];
  B.nDeep = B.p0 - B.r0, B.nPad = blockPadEnd[blk] - B.p0;
  B.steps = withResidual ? max(B.nDeep + 4, B.nPad + 6) : B.nDeep + 4;
  return B;
}
// The line wave gw (0..3 in its group) works on at local step ls of block B, or -1; kind 1 pass, 2 residual.
__device__ __forceinline__ int LineOfStep(const BandBlock& B, int gw, int ls, int withResidual, int& kind) {
  kind = 0;
  if (gw < 3) {
    const int row = ls - 2 * gw;
    if (row < 0 || row >= B.nDeep) return -1;
    kind = 1;
    return B.r0 + row;
  }
  const int pr = ls - 6;
  if (!withResidual || pr < 0 || pr >= B.nPad) return -1;
  kind = 2;
  return B.p0 + pr;
}
// A lane's elements of a line from the row element table (no search, no load that waits on another):
// a pass takes elements 2 lane and 2 lane + 1, a residual row lane, lane + 64 and lane + 128.
__device__ __forceinline__ RowLoc LoadRow(const int4* __restrict__ elems, int o0, int o1, int kind) {
  const int lane = threadIdx.x % kWave;
  RowLoc L;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int e = kind == 1 ? 2 * lane + q : lane + kWave * q;
    const bool in = kind != 0 && (kind == 2 || q < 2) && o0 + e < o1;
    const int4 v = elems[in ? o0 + e : 0];  // unconditional (element 0 where not in), then selected: no branch
    L.g[q] = in ? v.x : -1, L.p[q] = v.y, L.n[q] = v.z, L.d[q] = v.w >= 0 ? v.w : -v.w - 1, L.first[q] = v.w < 0;
  }
  return L;
}
// Where a group's blocks stand at a task step: the block it works on (or past the task's end) and
// the task step that block starts at.  G = 1: the blocks one after another; G > 1: every G-th block
// of the task, at the host schedule's blockStart.
template <int G>
struct BandCursor {
  int blk, b1, start;
  BandBlock B;
  __device__ __forceinline__ void Init(int first, int last, const int* __restrict__ blockBegin, const int* __restrict__ blockDeepEnd,
                                       const int* __restrict__ blockPadEnd, const int* __restrict__ blockStart, int withResidual) {
    blk = first, b1 = last, start = 0;
    if (blk < b1) {
      B = LoadBlock(blockBegin, blockDeepEnd, blockPadEnd, blk, withResidual);
      if (G > 1) start = blockStart[blk];
    }
  }
  // move to task step st (non-decreasing); returns the line of wave gw there, or -1
  __device__ __forceinline__ int At(int st, int gw, const int* __restrict__ blockBegin, const int* __restrict__ blockDeepEnd,
                                    const int* __restrict__ blockPadEnd, const int* __restrict__ blockStart, int withResidual, int& kind) {
    while (blk < b1 && st >= start + B.steps) {
      const int end = start + B.steps;
      blk += G;
      if (blk < b1) B = LoadBlock(blockBegin, blockDeepEnd, blockPadEnd, blk, withResidual), start = G == 1 ? end : blockStart[blk];
    }
    kind = 0;
    if (blk >= b1 || st < start) return -1;
    return LineOfStep(B, gw, st - start, withResidual, kind);
  }
};
template <int G, int KC>
__global__ __launch_bounds__(G * 4 * kWave) void KBandGs(int color, int nTasks, const int* __restrict__ taskPtr,
                                                         const int* __restrict__ blockBegin, const int* __restrict__ blockDeepEnd,
                                                         const int* __restrict__ blockPadEnd, const int* __restrict__ lineElem,
                                                         const int4* __restrict__ elems, const int* __restrict__ blockStart,
                                                         const int* __restrict__ taskSteps, const double* __restrict__ coef, int D,
                                                         double* x, const double* __restrict__ b, double* r, int K, int ld,
                                                         int withResidual) {
  const int t = 2 * blockIdx.x + color;
  if (t >= nTasks) return;
  const int wave = threadIdx.x / kWave, group = wave / 4, gw = wave % 4;
  const int b0 = taskPtr[t], b1 = taskPtr[t + 1];
  int total = 0;  // task steps
  if constexpr (G == 1) {
    for (int k = b0; k < b1; k++) total += LoadBlock(blockBegin, blockDeepEnd, blockPadEnd, k, withResidual).steps;
  } else {
    total = taskSteps[t];
  }
  // Two steps of lookahead per wave: the line of step st + 2 (its element range, scalar loads) and
  // the elements of step st + 1 (vector loads) are fetched while step st computes, so a step waits
  // only for its own stencil and x.
  BandCursor<G> ahead;
  ahead.Init(b0 + group, b1, blockBegin, blockDeepEnd, blockPadEnd, blockStart, withResidual);
  int kind0, kind1;
  const int l0 = ahead.At(0, gw, blockBegin, blockDeepEnd, blockPadEnd, blockStart, withResidual, kind0);
  const int l1 = ahead.At(1, gw, blockBegin, blockDeepEnd, blockPadEnd, blockStart, withResidual, kind1);
  RowLoc cur = LoadRow(elems, kind0 ? lineElem[l0] : 0, kind0 ? lineElem[l0 + 1] : 0, kind0);
  int curKind = kind0, nextKind = kind1;
  int n0 = kind1 ? lineElem[l1] : 0, n1 = kind1 ? lineElem[l1 + 1] : 0;
  for (int st = 0; st < total; st++) {
    int kind2;
    const int l2 = ahead.At(st + 2, gw, blockBegin, blockDeepEnd, blockPadEnd, blockStart, withResidual, kind2);
    const int m0 = kind2 ? lineElem[l2] : 0, m1 = kind2 ? lineElem[l2 + 1] : 0;
    const RowLoc nxt = LoadRow(elems, n0, n1, nextKind);
    if (curKind == 1) GsRow<KC>(cur, coef, D, x, b, K, ld);
    else if (curKind == 2) ResRow(cur, coef, D, x, b, r, K, ld);
    __syncthreads();
    cur = nxt, curKind = nextKind;
    nextKind = kind2, n0 = m0, n1 = m1;
  }
}

// ------------------------------------------------------------------ boundary rows
// vb = b_B - A_BD x   (IterativeSolvers.inl:410-419)
__global__ void KBoundaryRhs(int B, const int* __restrict__ bgi, const int* __restrict__ rp, const int* __restrict__ col,
                             const double* __restrict__ val, const double* __restrict__ x, const double* __restrict__ b, double* vb, int K,
                             int ld) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)B * K) return;
  int i = idx % B, c = idx / B;
  const double* X = x + (size_t)c * ld;
  double s = 0;
  for (int k = rp[i]; k < rp[i + 1]; k++) s += X[col[k]] * val[k];
  vb[(size_t)c * B + i] = b[(size_t)c * ld + bgi[i]] - s;
}

// r_B = b_B - A_BD x - A_BB x_B   (IterativeSolvers.inl:549-555)
__global__ void KBoundaryResidual(int B, const int* __restrict__ bgi, const int* __restrict__ rpd, const int* __restrict__ cold,
                                  const double* __restrict__ vald, const int* __restrict__ rpb, const int* __restrict__ colb,
                                  const double* __restrict__ valb, const double* __restrict__ x, const double* __restrict__ b, double* r,
                                  int K, int ld, int ldr) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)B * K) return;
  int i = idx % B, c = idx / B;
  const double* X = x + (size_t)c * ld;
  double s = 0;
  for (int k = rpd[i]; k < rpd[i + 1]; k++) s += X[cold[k]] * vald[k];
  double v = b[(size_t)c * ld + bgi[i]] - s;
  double t = 0;
  for (int k = rpb[i]; k < rpb[i + 1]; k++) t += X[bgi[colb[k]]] * valb[k];
  r[(size_t)c * ldr + bgi[i]] = v - t;
}

// ------------------------------------------------------------------ exact LDL^T solves
__global__ void KLdltGather(int n, const int* __restrict__ perm, const double* __restrict__ src, int lds, double* y, int K) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * K) return;
  int k = idx % n, c = idx / n;
  y[(size_t)c * n + k] = src[(size_t)c * lds + perm[k]];
}

__global__ void KLdltScatter(int n, const int* __restrict__ perm, const int* __restrict__ map, const double* __restrict__ y, double* dst,
                             int ldd, int K) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * K) return;
  int k = idx % n, c = idx / n;
  int t = perm[k];
  if (map) t = map[t];
  dst[(size_t)c * ldd + t] = y[(size_t)c * n + k];
}

__global__ void KLdltDiag(int n, const double* __restrict__ dinv, double* y, int K) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * K) return;
  y[idx] *= dinv[idx % n];
}

// Teams: one workgroup per elimination subtree, its dependency levels separated by barriers.
__global__ __launch_bounds__(256) void KLdltTeams(const int* __restrict__ teamLev, const int* __restrict__ levPtr,
                                                  const int* __restrict__ rows, const int* __restrict__ rp, const int* __restrict__ cl,
                                                  const double* __restrict__ vl, double* y, int n, int K) {
  const int t = blockIdx.x;
  for (int lv = teamLev[t]; lv < teamLev[t + 1]; lv++) {
    const int r0 = levPtr[lv], cnt = levPtr[lv + 1] - r0;
    for (int q = threadIdx.x; q < cnt * K; q += blockDim.x) {
      int row = rows[r0 + q % cnt], c = q / cnt;
      double* Y = y + (size_t)c * n;
      double s = Y[row];
      for (int k = rp[row]; k < rp[row + 1]; k++) s -= vl[k] * Y[cl[k]];
      Y[row] = s;
    }
    __syncthreads();
  }
}

// Top of the elimination tree: one workgroup, a wavefront per (row, column) dot product.
__global__ __launch_bounds__(1024) void KLdltTop(int lev0, int lev1, const int* __restrict__ levPtr, const int* __restrict__ rows,
                                                 const int* __restrict__ rp, const int* __restrict__ cl, const double* __restrict__ vl, double* y,
                                                 int n, int K) {
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave, nw = blockDim.x / kWave;
  for (int lv = lev0; lv < lev1; lv++) {
    const int r0 = levPtr[lv], cnt = levPtr[lv + 1] - r0;
    for (int q = wv; q < cnt * K; q += nw) {
      int row = rows[r0 + q % cnt], c = q / cnt;
      double* Y = y + (size_t)c * n;
      double s = 0;
      for (int k = rp[row] + lane; k < rp[row + 1]; k += kWave) s += vl[k] * Y[cl[k]];
      for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
      if (lane == 0) Y[row] -= s;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ transfers
// coarse boundary rhs = R_B r   (IterativeSolvers.inl:1027-1030)
__global__ void KCsrGatherScatter(int rows, const int* __restrict__ outMap, const int* __restrict__ rp, const int* __restrict__ cl,
                                  const double* __restrict__ vl, const double* __restrict__ in, int ldi, double* out, int ldo, int K) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)rows * K) return;
  int i = idx % rows, c = idx / rows;
  const double* X = in + (size_t)c * ldi;
  double s = 0;
  for (int k = rp[i]; k < rp[i + 1]; k++) s += X[cl[k]] * vl[k];
  out[(size_t)c * ldo + outMap[i]] = s;
}

// coarse deep rhs: 3x3 full weighting around the fine texel 2c (IterativeSolvers.inl:1034-1050)
__global__ void KRestrictDeep(int nLines, const int* __restrict__ lines, const double* __restrict__ in, int ldi, double* out, int ldo, int K) {
  const int l = blockIdx.x;
  if (l >= nLines) return;
  const int* L = lines + 5 * l;
  const int start = L[0], end = L[1], prev = L[2], next = L[3], outStart = L[4];
  const int len = (end - start) / 2 + 1;
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    const int j = 2 * i;
    for (int c = 0; c < K; c++) {
      const double* cu = in + (size_t)c * ldi + start;
      const double* pv = in + (size_t)c * ldi + prev;
      const double* nx = in + (size_t)c * ldi + next;
      out[(size_t)c * ldo + outStart + i] = cu[j] + (pv[j] + cu[j - 1] + cu[j + 1] + nx[j]) * 0.5 +
                                            (pv[j - 1] + pv[j + 1] + nx[j - 1] + nx[j + 1]) * 0.25;
    }
  }
}

// x_fine += P x_coarse, bilinear along prolongation lines (IterativeSolvers.inl:1122-1168)
__global__ void KProlongAdd(int nLines, const int* __restrict__ lines, const double* __restrict__ in, int ldi, double* out, int ldo, int K) {
  const int l = blockIdx.x;
  if (l >= nLines) return;
  const int* L = lines + 6 * l;
  const int start = L[0], len = L[1], center = L[2], prev = L[3], next = L[4], off = L[5] ? 0 : 1;
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    const int a = (i + off) / 2, bq = (i + off + 1) / 2;
    for (int c = 0; c < K; c++) {
      const double* X = in + (size_t)c * ldi;
      double v = center != -1 ? (X[center + a] + X[center + bq]) * 0.5 : (X[prev + a] + X[prev + bq] + X[next + a] + X[next + bq]) * 0.25;
      out[(size_t)c * ldo + start + i] += v;
    }
  }
}

// ------------------------------------------------------------------ operator application
// Deep rows of y = A x with raw stencils (MultiplyBySystemMatrix_NoReciprocals, IterativeSolvers.inl:881-950);
// with `b` given, writes the residual b - A x instead.
__global__ void KApplyDeep(int nLines, const int* __restrict__ lines, const double* __restrict__ coef, int D, const double* __restrict__ x,
                           const double* __restrict__ b, double* y, int K, int ld) {
  const int l = blockIdx.x;
  if (l >= nLines) return;
  const int* L = lines + 5 * l;
  const int start = L[0], len = L[1] - L[0] + 1, prev = L[2], next = L[3], d0 = L[4];
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    double c[9];
    for (int k = 0; k < 9; k++) c[k] = coef[(size_t)k * D + d0 + i];
    for (int cc = 0; cc < K; cc++) {
      const double* X = x + (size_t)cc * ld;
      double s = X[prev + i - 1] * c[0] + X[prev + i] * c[1] + X[prev + i + 1] * c[2] + X[start + i - 1] * c[3] + X[start + i] * c[4] +
                 X[start + i + 1] * c[5] + X[next + i - 1] * c[6] + X[next + i] * c[7] + X[next + i + 1] * c[8];
      y[(size_t)cc * ld + start + i] = b ? b[(size_t)cc * ld + start + i] - s : s;
    }
  }
}

__global__ void KApplyBoundary(int B, const int* __restrict__ bgi, const int* __restrict__ rpd, const int* __restrict__ cold,
                               const double* __restrict__ vald, const int* __restrict__ rpb, const int* __restrict__ colb,
                               const double* __restrict__ valb, const double* __restrict__ x, const double* __restrict__ b, double* y, int K,
                               int ld) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)B * K) return;
  int i = idx % B, c = idx / B;
  const double* X = x + (size_t)c * ld;
  double t = 0;
  for (int k = rpb[i]; k < rpb[i + 1]; k++) t += X[bgi[colb[k]]] * valb[k];
  double s = 0;
  for (int k = rpd[i]; k < rpd[i + 1]; k++) s += X[cold[k]] * vald[k];
  double v = t + s;
  y[(size_t)c * ld + bgi[i]] = b ? b[(size_t)c * ld + bgi[i]] - v : v;
}

// Squared 2-norm of column blockIdx.y: a bounded grid strides over the column so that only a
// few hundred blocks add into the column's accumulator.
__global__ __launch_bounds__(256) void KSquaredNorm(int N, const double* __restrict__ v, int ld, double* out) {
  const int c = blockIdx.y;
  double a = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double t = v[(long)c * ld + i];
    a += t * t;
  }
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  __shared__ double part[4];
  if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = a;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out + c, part[0] + part[1] + part[2] + part[3]);
}

__global__ void KZero(long n, double* x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// ------------------------------------------------------------------ host side
static inline unsigned Blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

DeviceSystem::~DeviceSystem() {
  for (void* p : allocations) (void)hipFree(p);
}

// ------------------------------------------------------------------ normalised-gradient integration
// One thread per cell sample: the bilinear gradient at the sample from its 4 corners, normalised in
// the metric, dotted with the 4 integrated basis gradients and added to the corners' rhs.
__global__ void KVfCell(long n, const int* __restrict__ corners, const float* __restrict__ data, double invW, double invH,
                        const double* __restrict__ u, int ldu, double* rhs, int ldr, int K) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int cn[4];
  for (int k = 0; k < 4; k++) cn[k] = corners[4 * s + k];
  const double x = data[s], y = data[n + s], t00 = data[2 * n + s], t01 = data[3 * n + s], t11 = data[4 * n + s];
  double vx[4], vy[4];
  for (int k = 0; k < 4; k++) vx[k] = data[(5 + 2 * k) * n + s], vy[k] = data[(6 + 2 * k) * n + s];
  for (int ch = 0; ch < K; ch++) {
    const double* U = u + (long)ch * ldu;
    const double c0 = U[cn[0]], c1 = U[cn[1]], c2 = U[cn[2]], c3 = U[cn[3]];
    // the bilinear gradient in difference form: exactly zero for a constant potential
    const double nx = -((1 - y) * (c1 - c0) + y * (c2 - c3)) * invW;
    const double ny = -((1 - x) * (c3 - c0) + x * (c2 - c1)) * invH;
    const double ax = t00 * nx + t01 * ny, ay = t01 * nx + t11 * ny;
    const double len2 = nx * ax + ny * ay;
    if (!(len2 > 0)) continue;
    const double il = 1.0 / sqrt(len2);
    double* R = rhs + (long)ch * ldr;
    for (int k = 0; k < 4; k++) atomicAdd(&R[cn[k]], (ax * vx[k] + ay * vy[k]) * il);
  }
}

// One thread per general sample (quadratic seam elements): gradient from its texel entries.
__global__ void KVfGeneral(int n, const int* __restrict__ ptr, const int* __restrict__ texel, const float* __restrict__ G,
                           const float* __restrict__ V, const float* __restrict__ T, const double* __restrict__ u, int ldu, double* rhs,
                           int ldr, int K) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int e0 = ptr[s], e1 = ptr[s + 1];
  const double t00 = T[3 * s], t01 = T[3 * s + 1], t11 = T[3 * s + 2];
  for (int ch = 0; ch < K; ch++) {
    const double* U = u + (long)ch * ldu;
    // the entries' G sum to zero: differences to the first entry's value keep a constant exact
    const double u0 = U[texel[e0]];
    double nx = 0, ny = 0;
    for (int e = e0 + 1; e < e1; e++) {
      const double du = U[texel[e]] - u0;
      nx -= G[2 * e] * du, ny -= G[2 * e + 1] * du;
    }
    const double ax = t00 * nx + t01 * ny, ay = t01 * nx + t11 * ny;
    const double len2 = nx * ax + ny * ay;
    if (!(len2 > 0)) continue;
    const double il = 1.0 / sqrt(len2);
    double* R = rhs + (long)ch * ldr;
    for (int e = e0; e < e1; e++) atomicAdd(&R[texel[e]], (ax * V[2 * e] + ay * V[2 * e + 1]) * il);
  }
}

__global__ void KZeroStrided(int N, double* x, int ld, int K) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * K) return;
  x[(t / N) * ld + t % N] = 0.0;
}

// ------------------------------------------------------------------ dense top of an LDL^T solve
// buf[c][a] = y[c][top[a]] - sum over team columns of L[top[a]][i] y[c][i]   (one wave per top row)
__global__ __launch_bounds__(256) void KTopCouple(int st, const int* __restrict__ top, const int* __restrict__ rp,
                                                  const int* __restrict__ cl, const double* __restrict__ vl, const double* __restrict__ y,
                                                  int n, double* buf, int K) {
  const int a = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (a >= st) return;
  for (int c = 0; c < K; c++) {
    const double* Y = y + (long)c * n;
    double s = 0.0;
    for (int p = rp[a] + lane; p < rp[a + 1]; p += kWave) s += vl[p] * Y[cl[p]];
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (lane == 0) buf[(long)c * st + a] = Y[top[a]] - s;
  }
}
// buf[c][a] = y[c][top[a]]
__global__ void KTopGather(int st, const int* __restrict__ top, const double* __restrict__ y, int n, double* buf, int K) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)st * K) return;
  buf[t] = y[(t / st) * n + top[t % st]];
}
// y[c][top[a]] = sum_b M[a][b] buf[c][b]   (one wave per row; M row-major st x st)
__global__ __launch_bounds__(256) void KTopDense(int st, const int* __restrict__ top, const double* __restrict__ M,
                                                 const double* __restrict__ buf, double* y, int n, int K) {
  const int a = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (a >= st) return;
  const double* row = M + (long)a * st;
  for (int c = 0; c < K; c++) {
    const double* in = buf + (long)c * st;
    double s = 0.0;
    for (int b = lane; b < st; b += kWave) s += row[b] * in[b];
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (lane == 0) y[(long)c * n + top[a]] = s;
  }
}

static thread_local size_t g_uploadBytes = 0;
size_t UploadByteCounter() { return g_uploadBytes; }

template <class T>
static T* UploadTo(std::vector<void*>& allocs, const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* p = nullptr;
  GDTP_HIP(hipMalloc(&p, v.size() * sizeof(T)));
  allocs.push_back(p);
  g_uploadBytes += v.size() * sizeof(T);
  GDTP_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

// Wave groups of a level's band launches: kBandGroups (blocks of a task overlapped in time) where a
// launch has no more tasks than the device has CUs (its 12-wave workgroups then run in one round)
// and the tasks have two or more blocks on average (else there is little to overlap); else 1.
// (C2 4096^2: levels 0 and 1 have 640 and 320 tasks per launch, level 2 160 tasks of 2.2 blocks.)
static int BandGroups(const BandTasks& T) {
  const int nt = (int)T.taskPtr.size() - 1;
  if (nt <= 0) return 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
  const double blocksPerTask = double(T.blockBegin.size()) / nt;
  return (nt + 1) / 2 <= cus && blocksPerTask >= 2.0 ? kBandGroups : 1;
}

template <class T>
static T* Upload(DeviceSystem& ds, const std::vector<T>& v) {
  return UploadTo(ds.allocations, v);
}

static double* Alloc(DeviceSystem& ds, size_t n) {
  double* p = nullptr;
  GDTP_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)));
  g_uploadBytes += std::max<size_t>(n, 1) * sizeof(double);
  ds.allocations.push_back(p);
  GDTP_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(double)));
  return p;
}

static DevCsr UploadCsr(DeviceSystem& ds, const Csr& m) {
  DevCsr d;
  d.rows = m.rows, d.nnz = (int)m.col.size();
  d.rowptr = Upload(ds, m.rowptr);
  d.col = Upload(ds, m.col);
  d.val = Upload(ds, m.val);
  return d;
}

static std::vector<double> SoA10(const std::vector<double>& aos) {
  size_t D = aos.size() / 10;
  std::vector<double> s(aos.size());
  for (size_t d = 0; d < D; d++)
    for (int k = 0; k < 10; k++) s[k * D + d] = aos[10 * d + k];
  return s;
}

// Level-0 deep stencils in symmetric five-word form (the mixed path's FP64 residual): slot k < 4 of
// texel g holds its coupling to its lower neighbour k (k = 3(dj+1) + (di+1): (-1,-1), (0,-1),
// (+1,-1), (-1,0)), slot 4 the diagonal, SoA [5][N] over global indices.  A deep row reads its upper
// coupling k = 5..8 from slot 8-k of the upper neighbour (g+1, next-1, next, next+1); a neighbour
// that is not deep gets that slot from the deep row's own stencil.  Assembled couplings agree with
// their mirror to rounding (within 1 ulp), so this is A up to 1 ulp in some off-diagonal entries.
static std::vector<double> SymStencil(const std::vector<double>& aos, const std::vector<Line5>& lines, int N) {
  std::vector<double> s(5 * (size_t)N, 0.0);
  std::vector<char> deep(N, 0);
  const long nl = (long)lines.size();
  bool bad = false;
#pragma omp parallel for schedule(dynamic, 256) reduction(|| : bad)
  for (long r = 0; r < nl; r++) {
    const Line5& x = lines[r];
    if (x.start < 0 || x.end >= N) bad = true;
    else
      for (int g = x.start; g <= x.end; g++) deep[g] = 1;
  }
  // every (slot, texel) has one writer: a deep texel its own slots, a non-deep texel's slot 8-k the
  // deep row below it in direction k
#pragma omp parallel for schedule(dynamic, 256) reduction(|| : bad)
  for (long r = 0; r < nl; r++) {
    const Line5& x = lines[r];
    for (int i = 0; i <= x.end - x.start && !bad; i++) {
      const size_t d = (size_t)x.coeff + i;
      const int cur = x.start + i, next = x.next + i;
      if (cur + 1 >= N || next - 1 < 0 || next + 1 >= N) { bad = true; break; }
      for (int k = 0; k < 5; k++) s[k * (size_t)N + cur] = aos[10 * d + k];
      const int up[4] = {cur + 1, next - 1, next, next + 1};
      for (int k = 5; k < 9; k++)
        if (!deep[up[k - 5]]) s[(8 - k) * (size_t)N + up[k - 5]] = aos[10 * d + k];
    }
  }
  if (bad) throw std::runtime_error("deep raster line outside the level or without its upper neighbours");
  return s;
}

// Partition the elimination tree into subtree teams (at most `smax` rows) plus the top part, and
// group rows by dependency level inside each team and inside the top.
void UploadLdlt(std::vector<void*>& ds, const Ldlt& F, DevLdlt& d, int kcap) {
  const int n = F.n;
  d.n = n;
  d.perm = UploadTo(ds, F.perm);
  d.lrRowptr = UploadTo(ds, F.Lrows.rowptr), d.lrCol = UploadTo(ds, F.Lrows.col), d.lrVal = UploadTo(ds, F.Lrows.val);
  d.lcRowptr = UploadTo(ds, F.Lcols.rowptr), d.lcCol = UploadTo(ds, F.Lcols.col), d.lcVal = UploadTo(ds, F.Lcols.val);
  std::vector<double> dinv(n);
  for (int i = 0; i < n; i++) dinv[i] = 1.0 / F.D[i];
  d.dinv = UploadTo(ds, dinv);
  std::vector<int> parent(n, -1), size(n, 1), team(n, -1);
  for (int i = 0; i < n; i++)
    if (F.Lcols.rowptr[i + 1] > F.Lcols.rowptr[i]) parent[i] = F.Lcols.col[F.Lcols.rowptr[i]];
  for (int i = 0; i < n; i++)
    if (parent[i] >= 0) size[parent[i]] += size[i];
  const int smax = std::max(64, n / 256);
  int nTeams = 0;
  for (int v = n - 1; v >= 0; v--) {
    if (size[v] > smax) continue;
    if (parent[v] < 0 || size[parent[v]] > smax) team[v] = nTeams++;
    else team[v] = team[parent[v]];
  }
  // forward levels (dependencies within the same team, or within the top)
  std::vector<int> lev(n, 0);
  for (int k = 0; k < n; k++) {
    int l = 0;
    for (int p = F.Lrows.rowptr[k]; p < F.Lrows.rowptr[k + 1]; p++) {
      int i = F.Lrows.col[p];
      if (team[i] == team[k]) l = std::max(l, lev[i] + 1);
    }
    lev[k] = l;
  }
  auto build = [&](std::vector<int>& levPtr, std::vector<int>& rows, std::vector<int>& teamLev, int& top0, int& top1) {
    // group key (team, level); teams first in team order, then the top part
    int G = nTeams + 1;
    std::vector<int> maxLev(G, -1);
    for (int k = 0; k < n; k++) {
      int g = team[k] < 0 ? nTeams : team[k];
      maxLev[g] = std::max(maxLev[g], lev[k]);
    }
    std::vector<int> groupBase(G + 1, 0);
    for (int g = 0; g < G; g++) groupBase[g + 1] = groupBase[g] + maxLev[g] + 1;
    int nGroups = groupBase[G];
    std::vector<int> cnt(nGroups + 1, 0);
    for (int k = 0; k < n; k++) {
      int g = team[k] < 0 ? nTeams : team[k];
      cnt[groupBase[g] + lev[k] + 1]++;
    }
    for (int i = 0; i < nGroups; i++) cnt[i + 1] += cnt[i];
    levPtr = cnt;
    rows.assign(n, 0);
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < n; k++) {
      int g = team[k] < 0 ? nTeams : team[k];
      rows[fillp[groupBase[g] + lev[k]]++] = k;
    }
    teamLev.assign(groupBase.begin(), groupBase.begin() + nTeams + 1);
    top0 = groupBase[nTeams], top1 = groupBase[G];
    int mr = 0;
    for (int i = 0; i < groupBase[nTeams]; i++) mr = std::max(mr, levPtr[i + 1] - levPtr[i]);
    d.maxTeamRows = std::max(d.maxTeamRows, mr);
  };
  std::vector<int> fLevPtr, fRows, fTeamLev, bLevPtr, bRows, bTeamLev;
  build(fLevPtr, fRows, fTeamLev, d.topFwdLevBegin, d.topFwdLevEnd);
  // backward levels: dependencies on ancestors in the same team (the top is solved first)
  std::fill(lev.begin(), lev.end(), 0);
  for (int i = n - 1; i >= 0; i--) {
    int l = 0;
    for (int p = F.Lcols.rowptr[i]; p < F.Lcols.rowptr[i + 1]; p++) {
      int k = F.Lcols.col[p];
      if (team[k] == team[i]) l = std::max(l, lev[k] + 1);
    }
    lev[i] = l;
  }
  build(bLevPtr, bRows, bTeamLev, d.topBwdLevBegin, d.topBwdLevEnd);
  d.nTeams = nTeams;
  d.fwdLevPtr = UploadTo(ds, fLevPtr), d.fwdRows = UploadTo(ds, fRows), d.teamFwdLev = UploadTo(ds, fTeamLev);
  d.bwdLevPtr = UploadTo(ds, bLevPtr), d.bwdRows = UploadTo(ds, bRows), d.teamBwdLev = UploadTo(ds, bTeamLev);
  // dense top: the top rows form a unit-lower block L_TT (separators after fill, nearly dense);
  // its inverse turns the sequential top solve into two dense matrix-vector products
  std::vector<int> top, local(n, -1);
  for (int k = 0; k < n; k++)
    if (team[k] < 0) local[k] = (int)top.size(), top.push_back(k);
  const int st = (int)top.size();
  if (st == 0 || st > kDenseTopMax || kcap < 1) return;
  std::vector<int> crp(1, 0), ccl;
  std::vector<double> cvl, Ltt((size_t)st * st, 0.0);
  for (int a = 0; a < st; a++) {
    const int k = top[a];
    Ltt[(size_t)a * st + a] = 1.0;
    for (int p = F.Lrows.rowptr[k]; p < F.Lrows.rowptr[k + 1]; p++) {
      const int i = F.Lrows.col[p];
      if (local[i] >= 0) Ltt[(size_t)a * st + local[i]] = F.Lrows.val[p];
      else ccl.push_back(i), cvl.push_back(F.Lrows.val[p]);
    }
    crp.push_back((int)ccl.size());
  }
  std::vector<double> inv((size_t)st * st, 0.0), invT((size_t)st * st, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < st; j++) {  // column j of the inverse: forward substitution from e_j
    std::vector<double> z(st, 0.0);
    z[j] = 1.0;
    for (int a = j + 1; a < st; a++) {
      double v = 0.0;
      const double* row = &Ltt[(size_t)a * st];
      for (int b = j; b < a; b++) v -= row[b] * z[b];
      z[a] = v;
    }
    for (int a = j; a < st; a++) inv[(size_t)a * st + j] = z[a], invT[(size_t)j * st + a] = z[a];
  }
  d.nTop = st, d.kcap = kcap;
  d.topRows = UploadTo(ds, top);
  d.topCplRowptr = UploadTo(ds, crp), d.topCplCol = UploadTo(ds, ccl), d.topCplVal = UploadTo(ds, cvl);
  d.topInv = UploadTo(ds, inv), d.topInvT = UploadTo(ds, invT);
  d.topBuf = UploadTo(ds, std::vector<double>((size_t)st * kcap, 0.0));
}

void UploadSystem(const Hierarchy& h, const OperatorHierarchy& oh, const FineOperator& op, int kcap, DeviceSystem& ds) {
  const int nl = (int)h.levels.size();
  ds.levels = nl, ds.kcap = kcap;
  g_uploadBytes = 0;
  ds.L.assign(nl, DevLevel());
  for (int l = 0; l < nl; l++) {
    const Level& H = h.levels[l];
    const LevelOperator& O = oh.levels[l];
    DevLevel& d = ds.L[l];
    d.N = H.numTexels, d.D = H.numDeep, d.B = H.numBoundary;
    d.coef = Upload(ds, SoA10(O.deepRecip));
    d.coefRaw = Upload(ds, SoA10(O.deepRaw));
    if (l == 0) d.coefSym = Upload(ds, SymStencil(O.deepRaw, H.rasterLines, H.numTexels));
    d.bgi = Upload(ds, H.boundaryGlobalIndex);
    d.bd = UploadCsr(ds, O.bd);
    d.bb = UploadCsr(ds, O.bb);
    if (l + 1 < nl) d.brestr = UploadCsr(ds, H.boundaryRestriction);
    d.nTasks = (int)H.tasks.taskPtr.size() - 1;
    d.taskPtr = Upload(ds, H.tasks.taskPtr);
    d.blockBegin = Upload(ds, H.tasks.blockBegin);
    d.blockDeepEnd = Upload(ds, H.tasks.blockDeepEnd);
    d.blockPadEnd = Upload(ds, H.tasks.blockPadEnd);
    d.blockStart = Upload(ds, H.tasks.blockStart);
    d.taskSteps = Upload(ds, H.tasks.taskSteps);
    d.bandGroups = BandGroups(H.tasks);
    d.lineElem = Upload(ds, H.tasks.lineElem);
    d.elems = reinterpret_cast<int4*>(Upload(ds, H.tasks.elems));
    std::vector<int> rl;
    for (auto& x : H.rasterLines) rl.insert(rl.end(), {x.start, x.end, x.prev, x.next, x.coeff});
    d.nRaster = (int)H.rasterLines.size();
    d.raster = Upload(ds, rl);
    rl.clear();
    for (auto& x : H.restrictionLines) rl.insert(rl.end(), {x.start, x.end, x.prev, x.next, x.coeff});
    d.nRestr = (int)H.restrictionLines.size();
    d.restr = Upload(ds, rl);
    rl.clear();
    for (auto& x : H.prolongationLines) rl.insert(rl.end(), {x.start, x.length, x.center, x.prev, x.next, x.aligned});
    d.nProl = (int)H.prolongationLines.size();
    d.prol = Upload(ds, rl);
    if (l + 1 < nl) UploadLdlt(ds.allocations, oh.boundaryFactor[l], d.factor, kcap);
    else UploadLdlt(ds.allocations, oh.coarseFactor, d.factor, kcap);
    d.x = Alloc(ds, (size_t)d.N * kcap);
    d.b = Alloc(ds, (size_t)d.N * kcap);
    d.r = Alloc(ds, (size_t)d.N * kcap);
    d.vb = Alloc(ds, (size_t)std::max(d.B, 1) * kcap);
    d.y = Alloc(ds, (size_t)std::max(d.factor.n, 1) * kcap);
  }
  for (int which = 0; which < 2; which++) {
    DevLevel& d = which == 0 ? ds.mass0 : ds.stiff0;
    const DevLevel& l0 = ds.L[0];
    d.N = l0.N, d.D = l0.D, d.B = l0.B, d.bgi = l0.bgi, d.nRaster = l0.nRaster, d.raster = l0.raster;
    d.coefRaw = Upload(ds, SoA10(which == 0 ? op.deepMass : op.deepStiffness));
    d.coefSym = Upload(ds, SymStencil(which == 0 ? op.deepMass : op.deepStiffness, h.levels[0].rasterLines, l0.N));
    d.bd = UploadCsr(ds, which == 0 ? op.bdMass : op.bdStiffness);
    d.bb = UploadCsr(ds, which == 0 ? op.bbMass : op.bbStiffness);
  }
  ds.bytes = g_uploadBytes;
}

void LdltSolve(DevLdlt& F, const double* src, int lds, double* dst, int ldd, const int* map, double* y, int K, hipStream_t s) {
  const int n = F.n;
  if (n == 0) return;
  const long nk = (long)n * K;
  KLdltGather<<<Blocks(nk, 256), 256, 0, s>>>(n, F.perm, src, lds, y, K);
  if (F.nTeams) KLdltTeams<<<F.nTeams, 256, 0, s>>>(F.teamFwdLev, F.fwdLevPtr, F.fwdRows, F.lrRowptr, F.lrCol, F.lrVal, y, n, K);
  const bool dense = F.nTop > 0 && K <= F.kcap;
  const unsigned topBlocks = Blocks(F.nTop, 256 / kWave);
  if (dense) {
    KTopCouple<<<topBlocks, 256, 0, s>>>(F.nTop, F.topRows, F.topCplRowptr, F.topCplCol, F.topCplVal, y, n, F.topBuf, K);
    KTopDense<<<topBlocks, 256, 0, s>>>(F.nTop, F.topRows, F.topInv, F.topBuf, y, n, K);
  } else if (F.topFwdLevEnd > F.topFwdLevBegin)
    KLdltTop<<<1, 1024, 0, s>>>(F.topFwdLevBegin, F.topFwdLevEnd, F.fwdLevPtr, F.fwdRows, F.lrRowptr, F.lrCol, F.lrVal, y, n, K);
  KLdltDiag<<<Blocks(nk, 256), 256, 0, s>>>(n, F.dinv, y, K);
  if (dense) {
    KTopGather<<<Blocks((long)F.nTop * K, 256), 256, 0, s>>>(F.nTop, F.topRows, y, n, F.topBuf, K);
    KTopDense<<<topBlocks, 256, 0, s>>>(F.nTop, F.topRows, F.topInvT, F.topBuf, y, n, K);
  } else if (F.topBwdLevEnd > F.topBwdLevBegin)
    KLdltTop<<<1, 1024, 0, s>>>(F.topBwdLevBegin, F.topBwdLevEnd, F.bwdLevPtr, F.bwdRows, F.lcRowptr, F.lcCol, F.lcVal, y, n, K);
  if (F.nTeams) KLdltTeams<<<F.nTeams, 256, 0, s>>>(F.teamBwdLev, F.bwdLevPtr, F.bwdRows, F.lcRowptr, F.lcCol, F.lcVal, y, n, K);
  KLdltScatter<<<Blocks(nk, 256), 256, 0, s>>>(n, F.perm, map, y, dst, ldd, K);
}

static void CheckK(const DeviceSystem& ds, int K) {
  if (K < 1 || K > ds.kcap) throw std::runtime_error("column count outside [1, capacity]");
}

void BoundarySolveLevel(DeviceSystem& ds, int l, const double* rhs, double* sol, int K, int ld, hipStream_t s) {
  CheckK(ds, K);
  DevLevel& d = ds.L[l];
  if (l + 1 >= ds.levels) throw std::runtime_error("the coarsest level has no boundary solver");
  LdltSolve(d.factor, rhs, ld, sol, ld, nullptr, d.y, K, s);
}

void CoarseSolve(DeviceSystem& ds, const double* rhs, double* sol, int K, int ld, hipStream_t s) {
  CheckK(ds, K);
  DevLevel& d = ds.L[ds.levels - 1];
  LdltSolve(d.factor, rhs, ld, sol, ld, nullptr, d.y, K, s);
}

void BandSweepLevel(DeviceSystem& ds, int l, double* x, const double* b, double* r, int K, int ld, hipStream_t s) {
  CheckK(ds, K);
  DevLevel& d = ds.L[l];
  if (d.nTasks == 0) return;
  const int withResidual = r != nullptr ? 1 : 0;
#define GDTP_BAND(G, KC)                                                                                                       \
  KBandGs<G, KC><<<blocks, G * 4 * kWave, 0, s>>>(color, d.nTasks, d.taskPtr, d.blockBegin, d.blockDeepEnd, d.blockPadEnd,  \
                                                  d.lineElem, d.elems, d.blockStart, d.taskSteps, d.coef, d.D, x, b, r, K, ld,     \
                                                  withResidual)
  // column chunks of K up to three, else of two (a chunk of four needs 168-179 registers a lane:
  // occupancy 2 or spills)
#define GDTP_BAND_KC(G)             \
  switch (K >= 4 ? 2 : K) {         \
    case 1: GDTP_BAND(G, 1); break; \
    case 2: GDTP_BAND(G, 2); break; \
    default: GDTP_BAND(G, 3); break; \
  }
  for (int color = 0; color < 2; color++) {
    const int blocks = color == 0 ? (d.nTasks + 1) / 2 : d.nTasks / 2;
    if (blocks == 0) continue;
    if (d.bandGroups == kBandGroups) {
      GDTP_BAND_KC(kBandGroups)
    } else {
      GDTP_BAND_KC(1)
    }
  }
#undef GDTP_BAND_KC
#undef GDTP_BAND
}

void RelaxLevel(DeviceSystem& ds, int l, double* x, const double* b, double* r, int K, int ld, bool withResidual, hipStream_t s) {
  CheckK(ds, K);
  DevLevel& d = ds.L[l];
  if (d.B > 0) {
    const long bk = (long)d.B * K;
    KBoundaryRhs<<<Blocks(bk, 256), 256, 0, s>>>(d.B, d.bgi, d.bd.rowptr, d.bd.col, d.bd.val, x, b, d.vb, K, ld);
    LdltSolve(d.factor, d.vb, d.B, x, ld, d.bgi, d.y, K, s);
  }
  BandSweepLevel(ds, l, x, b, withResidual ? r : nullptr, K, ld, s);
  if (withResidual && d.B > 0) {
    const long bk = (long)d.B * K;
    KBoundaryResidual<<<Blocks(bk, 256), 256, 0, s>>>(d.B, d.bgi, d.bd.rowptr, d.bd.col, d.bd.val, d.bb.rowptr, d.bb.col, d.bb.val, x, b, r,
                                                       K, ld, ld);
  }
}

void RestrictLevel(DeviceSystem& ds, int l, const double* rF, double* bC, int K, int ldF, int ldC, hipStream_t s) {
  CheckK(ds, K);
  DevLevel &f = ds.L[l], &c = ds.L[l + 1];
  if (c.B > 0)
    KCsrGatherScatter<<<Blocks((long)c.B * K, 256), 256, 0, s>>>(c.B, c.bgi, f.brestr.rowptr, f.brestr.col, f.brestr.val, rF, ldF, bC, ldC, K);
  if (c.nRestr > 0) KRestrictDeep<<<c.nRestr, 128, 0, s>>>(c.nRestr, c.restr, rF, ldF, bC, ldC, K);
}

void ProlongAddLevel(DeviceSystem& ds, int l, const double* xC, double* xF, int K, int ldC, int ldF, hipStream_t s) {
  CheckK(ds, K);
  DevLevel& f = ds.L[l];
  if (f.nProl > 0) KProlongAdd<<<f.nProl, 128, 0, s>>>(f.nProl, f.prol, xC, ldC, xF, ldF, K);
}

void ApplyLevel(DeviceSystem& ds, int l, const double* x, double* y, int K, int ld, hipStream_t s) {
  ApplyOperator(ds, ds.L[l], x, y, K, ld, s);
}

void ApplyOperator(DeviceSystem& ds, DevLevel& d, const double* x, double* y, int K, int ld, hipStream_t s) {
  CheckK(ds, K);
  if (d.B > 0)
    KApplyBoundary<<<Blocks((long)d.B * K, 256), 256, 0, s>>>(d.B, d.bgi, d.bd.rowptr, d.bd.col, d.bd.val, d.bb.rowptr, d.bb.col, d.bb.val, x,
                                                              nullptr, y, K, ld);
  if (d.nRaster > 0) KApplyDeep<<<d.nRaster, 128, 0, s>>>(d.nRaster, d.raster, d.coefRaw, d.D, x, nullptr, y, K, ld);
}

void SquaredNorms(const double* v, int N, int K, int ld, double* dNorms, hipStream_t s) {
  KZero<<<1, 64, 0, s>>>(K, dNorms);
  KSquaredNorm<<<dim3(std::min(Blocks(N, 256), 512u), K), 256, 0, s>>>(N, v, ld, dNorms);
}

void ResidualLevel(DeviceSystem& ds, int l, const double* x, const double* b, double* r, int K, int ld, hipStream_t s) {
  CheckK(ds, K);
  DevLevel& d = ds.L[l];
  if (d.B > 0)
    KApplyBoundary<<<Blocks((long)d.B * K, 256), 256, 0, s>>>(d.B, d.bgi, d.bd.rowptr, d.bd.col, d.bd.val, d.bb.rowptr, d.bb.col, d.bb.val, x, b,
                                                              r, K, ld);
  if (d.nRaster > 0) KApplyDeep<<<d.nRaster, 128, 0, s>>>(d.nRaster, d.raster, d.coefRaw, d.D, x, b, r, K, ld);
}

void VCycleReference(DeviceSystem& ds, double* x, const double* b, int K, int ld, hipStream_t s) {
  CheckK(ds, K);
  const int nl = ds.levels;
  for (int l = 1; l < nl; l++) {
    long n = (long)ds.L[l].N * K;
    KZero<<<Blocks(n, 256), 256, 0, s>>>(n, ds.L[l].x);
  }
  auto X = [&](int l) { return l == 0 ? x : ds.L[l].x; };
  auto Bv = [&](int l) { return l == 0 ? const_cast<double*>(b) : ds.L[l].b; };
  auto LD = [&](int l) { return l == 0 ? ld : ds.L[l].N; };
  for (int l = 0; l + 1 < nl; l++) {
    RelaxLevel(ds, l, X(l), Bv(l), ds.L[l].r, K, LD(l), true, s);
    // the level residual lives in r with the level's own stride
    RestrictLevel(ds, l, ds.L[l].r, ds.L[l + 1].b, K, LD(l), ds.L[l + 1].N, s);
  }
  CoarseSolve(ds, ds.L[nl - 1].b, ds.L[nl - 1].x, K, ds.L[nl - 1].N, s);
  for (int l = nl - 2; l >= 0; l--) {
    ProlongAddLevel(ds, l, ds.L[l + 1].x, X(l), K, ds.L[l + 1].N, LD(l), s);
    RelaxLevel(ds, l, X(l), Bv(l), nullptr, K, LD(l), false, s);
  }
}

DevVectorField::~DevVectorField() {
  for (void* p : allocations) (void)hipFree(p);
}

void UploadVectorField(const VectorFieldSamples& vs, DevVectorField& d) {
  d.nCell = (long)vs.numCellSamples();
  d.nGen = (int)vs.numGeneralSamples();
  d.invCellW = vs.invCellW, d.invCellH = vs.invCellH;
  std::vector<float> soa(13 * (size_t)d.nCell);
  for (long s = 0; s < d.nCell; s++)
    for (int f = 0; f < 13; f++) soa[(size_t)f * d.nCell + s] = vs.cellData[13 * (size_t)s + f];
  d.cellCorners = UploadTo(d.allocations, vs.cellCorners);
  d.cellData = UploadTo(d.allocations, soa);
  d.genPtr = UploadTo(d.allocations, vs.genPtr), d.genTexel = UploadTo(d.allocations, vs.genTexel);
  d.genTensor = UploadTo(d.allocations, vs.genTensor), d.genG = UploadTo(d.allocations, vs.genG), d.genV = UploadTo(d.allocations, vs.genV);
}

void VectorFieldDivergence(const DevVectorField& d, const double* u, int ldu, double* rhs, int ldr, int N, int K, hipStream_t s) {
  KZeroStrided<<<(unsigned)(((long)N * K + 255) / 256), 256, 0, s>>>(N, rhs, ldr, K);
  if (d.nCell > 0)
    KVfCell<<<(unsigned)((d.nCell + 255) / 256), 256, 0, s>>>(d.nCell, d.cellCorners, d.cellData, d.invCellW, d.invCellH, u, ldu, rhs, ldr,
                                                              K);
  if (d.nGen > 0)
    KVfGeneral<<<(unsigned)((d.nGen + 255) / 256), 256, 0, s>>>(d.nGen, d.genPtr, d.genTexel, d.genG, d.genV, d.genTensor, u, ldu, rhs,
                                                                ldr, K);
}

// ------------------------------------------------------------------ modulated stiffness
// Deep texels: the four cell products weighted by the masks of the cells (i-1,j-1), (i,j-1),
// (i-1,j), (i,j), summed in that order (IterativeSolvers.inl:1184-1192).
__global__ __launch_bounds__(256) void KModulatedDeep(int D, const int* __restrict__ deepGlobal, const int* __restrict__ cells,
                                                      const double* __restrict__ coef, const double* __restrict__ mask, double* out, int N,
                                                      int K) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D) return;
  const int cp = cells[2 * d], cn = cells[2 * d + 1];
  const double m0 = mask[cp], m1 = mask[cp + 1], m2 = mask[cn], m3 = mask[cn + 1];
  const long g = deepGlobal[d];
  for (int k = 0; k < K; k++) {
    const double* c = coef + (long)k * 4 * D + 4L * d;
    out[(long)k * N + g] = c[0] * m0 + c[1] * m1 + c[2] * m2 + c[3] * m3;
  }
}

// Boundary texels: a sparse row over cells (IterativeSolvers.inl:1177-1181).
__global__ __launch_bounds__(256) void KModulatedBoundary(int B, const int* __restrict__ bGlobal, const int* __restrict__ rp,
                                                          const int* __restrict__ cell, const double* __restrict__ val, long nnz,
                                                          const double* __restrict__ mask, double* out, int N, int K) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const int e0 = rp[r], e1 = rp[r + 1];
  const long g = bGlobal[r];
  for (int k = 0; k < K; k++) {
    const double* v = val + (long)k * nnz;
    double s = 0.0;
    for (int e = e0; e < e1; e++) s += v[e] * mask[cell[e]];
    out[(long)k * N + g] = s;
  }
}

DevModulation::~DevModulation() {
  for (void* p : allocations) (void)hipFree(p);
}

void UploadModulation(const Level& L0, const CellStiffnessRhs& c, DevModulation& m) {
  m.K = c.K, m.N = L0.numTexels, m.D = L0.numDeep, m.B = L0.numBoundary, m.numCells = c.numCells;
  m.nnz = (long)c.bCell.size();
  const size_t before = g_uploadBytes;
  m.deep = UploadTo(m.allocations, c.deep);
  m.deepCells = UploadTo(m.allocations, c.deepCells);
  m.deepGlobal = UploadTo(m.allocations, L0.deepGlobalIndex);
  m.bRowptr = UploadTo(m.allocations, c.bRowptr), m.bCell = UploadTo(m.allocations, c.bCell);
  m.bGlobal = UploadTo(m.allocations, L0.boundaryGlobalIndex);
  m.bVal = UploadTo(m.allocations, c.bVal);
  m.bytes = g_uploadBytes - before;
}

void ModulatedStiffness(const DevModulation& m, const double* mask, double* out, hipStream_t s) {
  if (m.D > 0)
    KModulatedDeep<<<(unsigned)((m.D + 255) / 256), 256, 0, s>>>(m.D, m.deepGlobal, m.deepCells, m.deep, mask, out, m.N, m.K);
  if (m.B > 0)
    KModulatedBoundary<<<(unsigned)((m.B + 255) / 256), 256, 0, s>>>(m.B, m.bGlobal, m.bRowptr, m.bCell, m.bVal, m.nnz, mask, out, m.N,
                                                                     m.K);
}

// ------------------------------------------------------------------ residual norms
// ||b - A x||^2 and ||b||^2 per column on level 0 without storing the residual: a fixed grid of
// kNormParts blocks strides over the raster lines (deep rows, raw stencil) and the boundary rows,
// each block writes one partial per (column, quantity), and a second pass adds the partials of a
// column in a fixed order, so repeated calls give bitwise identical norms.
template <int KC>
__global__ __launch_bounds__(256) void KResidualNormParts(int nLines, const int* __restrict__ lines, const double* __restrict__ coef, int D,
                                                          int B, const int* __restrict__ bgi, const int* __restrict__ rpd,
                                                          const int* __restrict__ cold, const double* __restrict__ vald,
                                                          const int* __restrict__ rpb, const int* __restrict__ colb,
                                                          const double* __restrict__ valb, const double* __restrict__ x,
                                                          const double* __restrict__ b, int ld, int c0, double* parts) {
  double rr[KC], bn[KC];
#pragma unroll
  for (int c = 0; c < KC; c++) rr[c] = 0.0, bn[c] = 0.0;
  for (int l = blockIdx.x; l < nLines; l += gridDim.x) {
    const int* L = lines + 5 * l;
    const int start = L[0], len = L[1] - L[0] + 1, prev = L[2], next = L[3], d0 = L[4];
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
      double cf[9];
#pragma unroll
      for (int k = 0; k < 9; k++) cf[k] = coef[(size_t)k * D + d0 + i];
#pragma unroll
      for (int c = 0; c < KC; c++) {
        const double* X = x + (size_t)(c0 + c) * ld;
        const double s = X[prev + i - 1] * cf[0] + X[prev + i] * cf[1] + X[prev + i + 1] * cf[2] + X[start + i - 1] * cf[3] +
                         X[start + i] * cf[4] + X[start + i + 1] * cf[5] + X[next + i - 1] * cf[6] + X[next + i] * cf[7] +
                         X[next + i + 1] * cf[8];
        const double t = b[(size_t)(c0 + c) * ld + start + i], r = t - s;
        rr[c] += r * r, bn[c] += t * t;
      }
    }
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < B; i += gridDim.x * blockDim.x) {
    const int g = bgi[i];
#pragma unroll
    for (int c = 0; c < KC; c++) {
      const double* X = x + (size_t)(c0 + c) * ld;
      double t = 0.0, s = 0.0;
      for (int k = rpb[i]; k < rpb[i + 1]; k++) t += X[bgi[colb[k]]] * valb[k];
      for (int k = rpd[i]; k < rpd[i + 1]; k++) s += X[cold[k]] * vald[k];
      const double bv = b[(size_t)(c0 + c) * ld + g], r = bv - (t + s);
      rr[c] += r * r, bn[c] += bv * bv;
    }
  }
  __shared__ double part[2 * KC][4];
#pragma unroll
  for (int c = 0; c < KC; c++) {
    double a = rr[c], e = bn[c];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64), e += __shfl_xor(e, o, 64);
    if (threadIdx.x % 64 == 0) part[2 * c][threadIdx.x / 64] = a, part[2 * c + 1][threadIdx.x / 64] = e;
  }
  __syncthreads();
  if (threadIdx.x < 2 * KC) {
    const int c = threadIdx.x / 2, w = threadIdx.x % 2;
    const double* p = part[threadIdx.x];
    parts[((size_t)(c0 + c) * 2 + w) * gridDim.x + blockIdx.x] = (p[0] + p[1]) + (p[2] + p[3]);
  }
}

// out[q] = sum of the nParts partials of row q (q = 2 c + w), in a fixed order.
__global__ __launch_bounds__(256) void KNormPartsSum(const double* __restrict__ parts, int nParts, double* out) {
  const int q = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < nParts; i += blockDim.x) a += parts[(size_t)q * nParts + i];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  __shared__ double w[4];
  if (threadIdx.x % 64 == 0) w[threadIdx.x / 64] = a;
  __syncthreads();
  if (threadIdx.x == 0) out[q] = (w[0] + w[1]) + (w[2] + w[3]);
}

void ResidualNorms(DeviceSystem& ds, const double* x, const double* b, int K, int ld, double* parts, double* out, hipStream_t s) {
  DevLevel& d = ds.L[0];
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = std::min(4, K - c0);
#define GDTP_NORM_PARTS(KC)                                                                                                         \
  KResidualNormParts<KC><<<kNormParts, 256, 0, s>>>(d.nRaster, d.raster, d.coefRaw, d.D, d.B, d.bgi, d.bd.rowptr, d.bd.col, d.bd.val, \
                                                    d.bb.rowptr, d.bb.col, d.bb.val, x, b, ld, c0, parts)
    switch (kc) {
      case 4: GDTP_NORM_PARTS(4); break;
      case 3: GDTP_NORM_PARTS(3); break;
      case 2: GDTP_NORM_PARTS(2); break;
      default: GDTP_NORM_PARTS(1); break;
    }
#undef GDTP_NORM_PARTS
  }
  KNormPartsSum<<<2 * K, 256, 0, s>>>(parts, kNormParts, out);
}

// ------------------------------------------------------------------ residual norms without allocation
// For callers that cannot allocate device memory: the partial sums and squared norms of
// ResidualNorms live in this file-scope array (one copy per GPU, held until the process ends).
// Host code never passes `scratch` itself to a kernel: its device address, looked up once per
// device with hipGetSymbolAddress, is kept in g_scratchAddr.  Every call on one GPU shares the
// array, so calls that overlap on one GPU race; each call overwrites every entry it reads.
static __device__ __align__(256) double scratch[NormPartsSize(kNoAllocMaxCols) + 2 * kNoAllocMaxCols];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratchLock;
static double* g_scratchAddr[kMaxDevices] = {};

static double* ScratchAddress() {
  int dev = 0;
  GDTP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal beyond the no-allocation scratch table");
  std::lock_guard<std::mutex> lock(g_scratchLock);
  if (!g_scratchAddr[dev]) {
    void* p = nullptr;
    GDTP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratchAddr[dev] = static_cast<double*>(p);
  }
  return g_scratchAddr[dev];
}

void ResidualNormsNoAlloc(DeviceSystem& ds, const double* x, const double* b, int K, int ld, double* hostSquares, hipStream_t s) {
  if (K < 1 || K > kNoAllocMaxCols) throw std::invalid_argument("column count outside the no-allocation scratch");
  double* p = ScratchAddress();
  double* out = p + NormPartsSize(kNoAllocMaxCols);
  ResidualNorms(ds, x, b, K, ld, p, out, s);
  GDTP_HIP(hipMemcpyAsync(hostSquares, out, 2 * (size_t)K * sizeof(double), hipMemcpyDeviceToHost, s));
  GDTP_HIP(hipStreamSynchronize(s));
}

}  // namespace gdtp
