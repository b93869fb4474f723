// Device-resident multigrid system and the HIP V-cycle (internal interface of libgdtp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "gdtp_galerkin.h"
#include "gdtp_vectorfield.h"

namespace gdtp {

#define GDTP_HIP(call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Exact triangular solves of an LDL^T factor, split into independent elimination-subtree
// teams (one workgroup each) and a top part solved by a single workgroup.
struct DevLdlt {
  int n = 0;
  int *perm = nullptr;
  int *lrRowptr = nullptr, *lrCol = nullptr, *lcRowptr = nullptr, *lcCol = nullptr;
  double *lrVal = nullptr, *lcVal = nullptr, *dinv = nullptr;
  // teams: team t owns forward levels [teamLevPtr[t], teamLevPtr[t+1]) of levRowPtr -> levRows
  int nTeams = 0;
  int *teamFwdLev = nullptr, *teamBwdLev = nullptr;  // nTeams+1 each, indices into the level tables
  int *fwdLevPtr = nullptr, *fwdRows = nullptr, *bwdLevPtr = nullptr, *bwdRows = nullptr;
  int topFwdLevBegin = 0, topFwdLevEnd = 0, topBwdLevBegin = 0, topBwdLevEnd = 0;
  int maxTeamRows = 0;
  // dense top (when the top has at most kDenseTopMax rows): the top rows (permuted indices), their
  // coupling to team rows (CSR over the top rows), the inverse of the unit-lower top block and its
  // transpose (row-major s x s), and a [kcap][s] scratch vector
  int nTop = 0, kcap = 0;
  int *topRows = nullptr, *topCplRowptr = nullptr, *topCplCol = nullptr;
  double *topCplVal = nullptr, *topInv = nullptr, *topInvT = nullptr, *topBuf = nullptr;
};

struct DevCsr {
  int rows = 0, nnz = 0;
  int *rowptr = nullptr, *col = nullptr;
  double* val = nullptr;
};

struct DevLevel {
  int N = 0, D = 0, B = 0;
  double* coef = nullptr;      // reciprocal (levels < L-1) deep stencil, SoA [10][D]
  double* coefRaw = nullptr;   // raw deep stencil A, SoA [10][D]
  double* coefSym = nullptr;   // level 0: A's deep couplings in symmetric five-word form, SoA [5][N] over global
                               // indices (slots 0-3: lower neighbours k = 0..3, slot 4: diagonal)
  int* bgi = nullptr;          // boundary -> global
  DevCsr bd, bb, brestr;       // A_BD, A_BB, boundary restriction to level+1
  // band tasks
  int nTasks = 0;
  int *taskPtr = nullptr, *blockBegin = nullptr, *blockDeepEnd = nullptr, *blockPadEnd = nullptr;
  int *blockStart = nullptr, *taskSteps = nullptr;  // pipeline schedule: first task step of each block, steps per task
  int bandGroups = 1;                                // wave groups of the band launches (1 or kBandGroups)
  int* lineElem = nullptr;                           // row elements of every band line (offsets)
  int4* elems = nullptr;                             // (g, above, below, deep index or -d-1 at a segment start)
  int nRaster = 0, nRestr = 0, nProl = 0;
  int *raster = nullptr, *restr = nullptr, *prol = nullptr;  // 5, 5, 6 ints per line
  DevLdlt factor;              // boundary factor (levels < L-1) or full coarse factor (level L-1)
  // work vectors for up to `kcap` columns
  double *x = nullptr, *b = nullptr, *r = nullptr, *vb = nullptr, *y = nullptr;
};

struct DeviceSystem {
  int levels = 0, kcap = 0;
  std::vector<DevLevel> L;
  DevLevel mass0, stiff0;      // level-0 mass and stiffness (raw stencils and boundary rows only)
  std::vector<void*> allocations;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  ~DeviceSystem();
};

// Upload an LDL^T factor with its team / top level schedule; LdltSolve solves for K columns,
// gathering src[perm] and scattering into dst (through `map` when not null).
void UploadLdlt(std::vector<void*>& allocations, const Ldlt& F, DevLdlt& d, int kcap);
// Bytes uploaded / allocated by this thread's device uploads so far (monotone counter).
size_t UploadByteCounter();
void LdltSolve(DevLdlt& F, const double* src, int lds, double* dst, int ldd, const int* map, double* y, int K, hipStream_t s);

void UploadSystem(const Hierarchy& h, const OperatorHierarchy& oh, const FineOperator& op, int kcap, DeviceSystem& ds);

// One V-cycle of the reference schedule on level-0 vectors x (in/out) and b, K <= kcap columns,
// column stride ld (IterativeSolvers.inl:1211-1333).
void VCycleReference(DeviceSystem& ds, double* x, const double* b, int K, int ld, hipStream_t s);

// Building blocks, exported for kernel-level parity tests.
void RelaxLevel(DeviceSystem& ds, int level, double* x, const double* b, double* r, int K, int ld, bool withResidual, hipStream_t s);
void BandSweepLevel(DeviceSystem& ds, int level, double* x, const double* b, double* r, int K, int ld, hipStream_t s);
void RestrictLevel(DeviceSystem& ds, int level, const double* rFine, double* bCoarse, int K, int ldF, int ldC, hipStream_t s);
void ProlongAddLevel(DeviceSystem& ds, int level, const double* xCoarse, double* xFine, int K, int ldC, int ldF, hipStream_t s);
void BoundarySolveLevel(DeviceSystem& ds, int level, const double* rhs, double* sol, int K, int ld, hipStream_t s);
void CoarseSolve(DeviceSystem& ds, const double* rhs, double* sol, int K, int ld, hipStream_t s);
// y = A x with raw (non-reciprocal) coefficients of `level`; residual r = b - A x.
void ApplyLevel(DeviceSystem& ds, int level, const double* x, double* y, int K, int ld, hipStream_t s);
void ApplyOperator(DeviceSystem& ds, DevLevel& op, const double* x, double* y, int K, int ld, hipStream_t s);
// Squared 2-norms of K columns (stride ld) accumulated into dNorms[0..K) (zeroed first).
void SquaredNorms(const double* v, int N, int K, int ld, double* dNorms, hipStream_t s);
void ResidualLevel(DeviceSystem& ds, int level, const double* x, const double* b, double* r, int K, int ld, hipStream_t s);
// Squared norms of the level-0 residual b - A x and of b for K columns: out[2c] = ||b_c - A x_c||^2,
// out[2c + 1] = ||b_c||^2 (device, 2K doubles); `parts` is a device scratch of NormPartsSize(K)
// doubles, every entry of which the call overwrites.
constexpr int kNormParts = 512;
constexpr size_t NormPartsSize(int K) { return (size_t)2 * K * kNormParts; }
void ResidualNorms(DeviceSystem& ds, const double* x, const double* b, int K, int ld, double* parts, double* out, hipStream_t s);
// The same with the scratch in a static device array (no allocation): at most kNoAllocMaxCols
// columns; the 2 K squared norms are copied to hostSquares and the stream is synchronised.
constexpr int kNoAllocMaxCols = 64;
void ResidualNormsNoAlloc(DeviceSystem& ds, const double* x, const double* b, int K, int ld, double* hostSquares, hipStream_t s);

// Normalised-gradient integration samples on the device (cell samples SoA [13][n]).
struct DevVectorField {
  long nCell = 0;
  int nGen = 0;
  int* cellCorners = nullptr;  // 4 per sample
  float* cellData = nullptr;   // SoA [13][nCell]
  double invCellW = 0, invCellH = 0;
  int *genPtr = nullptr, *genTexel = nullptr;
  float *genTensor = nullptr, *genG = nullptr, *genV = nullptr;
  std::vector<void*> allocations;
  ~DevVectorField();
};
void UploadVectorField(const VectorFieldSamples& vs, DevVectorField& d);
// rhs[:, t] = sum over samples of < -grad u / |grad u|_g , v(t) > for K columns of level-0 vectors
// (IntergrateNormalizedVectorFieldFromSamples, VectorFieldIntergration.inl:452-582); rhs is
// overwritten.
void VectorFieldDivergence(const DevVectorField& d, const double* u, int ldu, double* rhs, int ldr, int N, int K, hipStream_t s);

// Per-cell modulated stiffness of fixed signal columns (the sharpening brush's right-hand side,
// CellStiffnessToTexelStiffness, IterativeSolvers.inl:1170-1208): t_k = sum_c mask_c S_c f_k.
struct DevModulation {
  int K = 0, N = 0, D = 0, B = 0, numCells = 0;
  long nnz = 0;
  double* deep = nullptr;       // [K][4D]: the four cells around each deep texel
  int* deepCells = nullptr;     // [2D]: cells (i-1,j-1) and (i-1,j) of deep texel d
  int* deepGlobal = nullptr;    // [D]
  int *bRowptr = nullptr, *bCell = nullptr, *bGlobal = nullptr;
  double* bVal = nullptr;       // [K][nnz]
  size_t bytes = 0;
  std::vector<void*> allocations;
  ~DevModulation();
};
void UploadModulation(const Level& L0, const CellStiffnessRhs& c, DevModulation& m);
// out[k][t] = sum over the cells c of texel t of mask[c] * (S_c f_k)_t for every level-0 texel t
// (out is [K][N], overwritten; mask has one weight per active cell).
void ModulatedStiffness(const DevModulation& m, const double* mask, double* out, hipStream_t s);

}  // namespace gdtp
