// Mixed-precision multi-RHS solver: FP64 residual and correction (iterative refinement) around
// an FP32 V-cycle on dense per-chart colour planes.
#pragma once
#include "gdtp_device.h"

namespace gdtp {

struct MixChart {  // per chart geometry of one level
  int w, h, w2, h2, base2, cofx, cofy;
};

// Tile shapes of the fused four-colour deep sweep: a tile of ta x tb sub-grid texels per
// workgroup of `threads` threads, and a variant.  The system uses one of them per level (or, with
// shape -1, one launch per colour on 64 x 4 tiles).
// Variants: the operands of a colour (stencil, rhs) are fetched right before it is computed
// (kFusedPerColour: fewest registers, most resident workgroups), one colour ahead
// (kFusedOneAhead), or all four colours' before the first is computed (kFusedAllAhead: one memory
// round trip per tile, most registers); kFusedXcd hands each of the eight XCDs (workgroup b runs on
// XCD b % 8) a contiguous eighth of the tile list, so that tiles sharing halo texels share an L2.
// A workgroup of `threads` threads covers a colour's region (the tile and the halo texels later
// colours read) in ceil(region / threads) passes: 384 threads cover every region of a 64 x 4
// tile in one pass, 640 those of a 64 x 8 tile.
constexpr int kFusedPerColour = 0, kFusedOneAhead = 1, kFusedAllAhead = 2, kFusedXcd = 4;
struct FusedShapeInfo {
  int ta, tb, threads, variant;
};
constexpr int kFusedShapes = 7;
constexpr FusedShapeInfo kFusedShape[kFusedShapes] = {
    {64, 4, 256, kFusedOneAhead},
    {64, 4, 256, kFusedOneAhead | kFusedXcd},
    {64, 4, 384, kFusedOneAhead | kFusedXcd},
    {64, 4, 384, kFusedAllAhead | kFusedXcd},
    {64, 8, 640, kFusedOneAhead | kFusedXcd},
    {64, 8, 640, kFusedAllAhead | kFusedXcd},
    {64, 4, 256, kFusedAllAhead | kFusedXcd},
};

// Deep sweeps that are not fused: one launch per colour on 64 x 4 tiles (kShapePerColour), or
// two launches per sweep, one per colour pair, with a workgroup per sub-grid row
// (kShapePairs; active row spans of up to kPairMaxSpan x kPairThreadsMax sub-grid texels).
constexpr int kShapePerColour = -1, kShapePairs = -4;
constexpr int kPairThreadsMax = 1024, kPairMaxSpan = 2;

struct DevMixLevel {
  int N = 0, B = 0, nCharts = 0;
  long plane = 0, Nd = 0;          // dense colour-plane size and total dense size (4 planes)
  MixChart* charts = nullptr;      // device
  std::vector<MixChart> hcharts;   // host copy
  float* stencil = nullptr;        // [10][Nd]: c0..c8 reciprocal form, slot 9 = a4; zero off deep texels (FP32 smoother only)
  uint4* stencilH = nullptr;       // [Nd]: the smoother's off-diagonal ratios c0..c3, c5..c8 as FP16 pairs
  float2* stencilC = nullptr;      // [Nd]: the smoother's c4 and the FP16 rounding sum sum_k (c_k - fp16(c_k))
  float* stencilSym = nullptr;     // [5][Nd]: raw A in symmetric five-word form for residuals (slots 0-3: lower
                                   // neighbours (-1,-1) (0,-1) (+1,-1) (-1,0); slot 4: a4, zero off deep texels)
  int8_t* type = nullptr;          // [Nd]: -1 inactive, 0/1 boundary, 2 deep
  int* g2d = nullptr;              // [N] compact global index -> dense index
  int nTiles = 0;
  bool chartSweeps = false;         // small charts: one workgroup per chart runs all four colours of a sweep
  int* tiles = nullptr;            // 3 ints per tile: chart, a0, b0 (sub-grid origin)
  unsigned long long* deepMask = nullptr;    // [tile][plane][tile row]: bit x set where texel (a0 + x, b0 + row) is deep
  unsigned long long* activeMask = nullptr;  // the same for active texels (deep or boundary)
  int* ftiles[kFusedShapes] = {};  // the same for each fused-sweep tile shape
  int nFTiles[kFusedShapes] = {};
  // colour-pair sweep: 4 ints (chart, b, first texel, span end) per wave of packed rows (pairSpan 1:
  // pairThreads / 64 per workgroup) or per row (wider spans: one workgroup each); nRows workgroups
  int* rows = nullptr;
  unsigned long long* rowMask = nullptr;  // packed rows: [wave][plane] deep texels of the wave's 64 texels (bit = lane)
  int nRows = 0, pairSpan = 0;     // sub-grid texels per thread of the widest span (1, 2; 0: too wide)
  int pairThreads = 0;             // threads per workgroup of the pair sweep (512, kPairThreadsMax)
  int shape = 0;                   // deep-sweep tile shape of this level (kFusedShape index, kShapePerColour, kShapePairs)
  // boundary rows: dense index, off-diagonal row (A_BB and A_BD merged, dense columns, one word per entry), diagonal
  int* bdIdx = nullptr;
  int* bRowptr = nullptr;
  int2* bEnt = nullptr;  // off-diagonal entries: (dense column, FP32 value bits)
  float* bDiag = nullptr;
  // overlapping Schwarz blocks of boundary rows: block b updates its core rows (dense indices
  // blkCore[blkCorePtr[b]..]) from the residual at its extended rows (boundary indices
  // blkExt[blkExtPtr[b]..]) through the core rows of the inverse of A restricted to the extended
  // rows, stored column-major at blkMat + blkMatPtr[b] ([ext][core]) in FP32, or at blkMatH +
  // blkMatPtrH[b] as FP16 quads ([ext / 4][core], extended rows 4s .. 4s + 3 in one 64-bit word,
  // lowest first) scaled per core row (entries |.| <= 1, core row r's scale blkScale[r])
  int nBlocks = 0;
  int *blkCorePtr = nullptr, *blkCore = nullptr, *blkExtPtr = nullptr, *blkExt = nullptr;
  long *blkMatPtr = nullptr, *blkMatPtrH = nullptr;
  float* blkMat = nullptr;
  uint2* blkMatH = nullptr;
  float* blkScale = nullptr;
  // restriction into this level (coarse boundary rows: fine dense columns, weights)
  int* rRowptr = nullptr;
  int2* rEnt = nullptr;  // (fine dense column, FP32 weight bits)
  // vectors [K][Nd] and boundary residual scratch [K][B]; the fused sweep reads e and writes e2
  // (levels swept by tiles), after which the two pointers swap
  float *e = nullptr, *rhs = nullptr, *res = nullptr, *tmpB = nullptr, *e2 = nullptr;
  float* eHome = nullptr;          // the buffer e was allocated as (level 0 returns to it between refinement steps)
};

struct MixedSystem {
  int levels = 0, kcap = 0;
  int nuPre = 1, nuPost = 2, boundarySweeps = 1;  // Schwarz sweeps over the boundary rows per Schwarz step
  bool fuseProlong = false;                       // fused levels: the first post sweep adds the prolongation itself
  float omega = 1.0f;                             // damping of the Schwarz block update
  std::vector<DevMixLevel> L;
  DevLdlt coarse;                  // exact FP64 solve of the coarsest level
  double *coarseRhs = nullptr, *coarseX = nullptr, *coarseY = nullptr;
  double* coarseInv = nullptr;     // dense inverse of a small coarsest level (row-major), else null
  double* norms = nullptr;         // device [2][kcap][64]: squared residual / rhs norm partial sums
  double* hostNorms = nullptr;     // pinned copy
  double* weights = nullptr;       // device [kcap][2]: per-column (mass, stiffness) weights of weighted solves
  double screen = 1.0, stiff = 0.0;  // the system's own weights
  // flexible conjugate gradients around the V-cycle (else plain iterative refinement)
  bool krylov = false;
  // the Schwarz steps read a level's FP16 block-inverse quads where it has them (set per solve:
  // plain refinement; flexible CG reads the FP32 inverses)
  bool schwarzHalf = true;
  // V-cycles per outer iteration (the level-0 correction carries over between them); 0 = automatic:
  // 2 for plain refinement unless the operator is nearly singular for FP32 (see rowSumRatio), else 1
  int cycles = 0;
  // median over level-0 deep rows of (row sum / diagonal) of the FP64 operator.  Continuing the FP32
  // correction through a second V-cycle evaluates A e in FP32, whose stencil rounding (~1e-7 of the
  // diagonal, coherent over regular regions) swamps the lowest eigenvalues once this ratio is tiny
  // (e.g. 1e-6 M + S, the geodesic distance system): the second cycle then amplifies the near-null
  // mode instead of reducing it.
  double rowSumRatio = 1.0;
  double *r64 = nullptr, *p64 = nullptr, *q64 = nullptr, *z64 = nullptr;  // [kcap][N0] each
  int* deepLine = nullptr;         // [D0]: raster line of each level-0 deep texel (FP64 outer kernels)
  double* dots = nullptr;          // [3][kcap][1024]: per-block partial sums of (z,q), (r,p), (p,q)
  double *alpha = nullptr, *beta = nullptr, *pqPrev = nullptr;  // [kcap]
  int* first = nullptr;            // first step of a solve (beta = 0)
  int* fresh = nullptr;            // x not yet written (refinement from a zero guess): the next update sets x = e
  std::vector<void*> allocations;
  size_t bytes = 0;                // device bytes held
  // captured refinement steps, one per (K, x, b, weighted) in use (least recently used evicted)
  struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int K = 0, cycles = 1;
    const void *x = nullptr, *b = nullptr;
    bool weighted = false;
    long lastUse = 0;
  };
  std::vector<GraphEntry> graphs;
  long graphClock = 0;
  hipStream_t stream = nullptr;
  ~MixedSystem();
};

void UploadMixed(const Hierarchy& h, const OperatorHierarchy& oh, DeviceSystem& ds, int kcap, MixedSystem& ms);

// Solve A x = b (level-0 compact FP64 vectors, K columns) until every column's relative residual
// is <= tol or maxIter refinement steps; x is used as the initial guess.  Returns the iterations
// run; relRes (host, K entries) receives the final relative residuals.  With `weights` (host,
// K pairs), column c solves (w[c][0] M + w[c][1] S) x = b instead, by flexible CG with the
// system's own V-cycle as the preconditioner of every column.
// `between` (may be null) is called with `ctx` after every iteration that leaves the solve
// unconverged, on the calling thread, with the device idle: a caller can yield the GPU there
// without ending the solve (flexible CG keeps its search direction).
// zeroGuess: start from x = 0 whatever x holds (x is zeroed, and the first residual is b itself).
int SolveMixed(DeviceSystem& ds, MixedSystem& ms, double* x, const double* b, int K, double tol, int maxIter, double* relRes,
               hipStream_t s, bool useGraph, const double* weights = nullptr, void (*between)(void*) = nullptr, void* ctx = nullptr,
               bool zeroGuess = false);

// Drop the captured refinement step (after the smoothing configuration changes).
void InvalidateGraph(MixedSystem& ms);

// One FP32 V-cycle of the correction equation on the current level-0 residual; zeroE0: the level-0
// correction is zero on entry (the first cycle after a residual was written).
void MixedVCycle(DeviceSystem& ds, MixedSystem& ms, int K, hipStream_t s, bool zeroE0);
// One four-colour deep Gauss-Seidel sweep of `level` on the solver's internal vectors (pre order);
// returns the number of kernel launches it made.
int MixedDeepSweep(MixedSystem& ms, int level, int K, hipStream_t s);

}  // namespace gdtp
