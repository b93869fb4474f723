/* gdtp: texel-domain screened-Poisson multigrid on MI355X (gfx950) — C ABI.
 *
 * Conventions
 *   - Every entry point returns GDTP_OK (0) or a nonzero status; gdtp_last_error() gives the
 *     message of the last failure on the calling thread.  (The reference returns int 1/0 and
 *     printf()s; include/Src/IterativeSolvers.inl:1211, UpdateCoefficients.inl:463.)
 *   - Texel vectors are column-major: `ncols` columns of N_level doubles, contiguous
 *     ([ncols][N]); device entry points take device pointers and an optional hipStream_t
 *     passed as void* (NULL = default stream).  Texels are in the reference's global raster order
 *     (charts in order, rows j ascending, columns i ascending; HierarchyConstruction.inl:259-290).
 *   - A handle is externally synchronised; distinct handles may be used from distinct threads.
 */
#ifndef GDTP_H
#define GDTP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GDTP_OK = 0,
  GDTP_ERROR_INVALID = 1,   /* bad argument or handle */
  GDTP_ERROR_GEOMETRY = 2,  /* mesh / atlas / hierarchy construction failed */
  GDTP_ERROR_NUMERIC = 3,   /* factorisation or solve failed (e.g. not positive definite) */
  GDTP_ERROR_DEVICE = 4     /* HIP runtime error or no device */
};

enum { GDTP_METRIC_EMBEDDING = 0, GDTP_METRIC_UNIFORM = 1 };
enum { GDTP_OPERATOR_SYSTEM = 0, GDTP_OPERATOR_MASS = 1, GDTP_OPERATOR_STIFFNESS = 2 };
enum { GDTP_SOLVE_FP64_REFERENCE = 0, GDTP_SOLVE_MIXED = 1, GDTP_SOLVE_ZERO_GUESS = 0x100 };
enum { GDTP_ASSEMBLY_DEVICE = 1 };
enum { GDTP_FUSE_PROLONGATION = 1 };
enum { GDTP_TILE_SHAPE_AUTO = -3, GDTP_TILE_SHAPE_PAIRS = -4 };

typedef struct gdtp_mesh gdtp_mesh;
typedef struct gdtp_hierarchy gdtp_hierarchy;
typedef struct gdtp_system gdtp_system;
typedef struct gdtp_modulation gdtp_modulation;

const char* gdtp_last_error(void);
/* Identifier of the sources this library was built from (a hash of the package's csrc/ files);
 * the Python binding refuses a library whose identifier differs from the sources beside it.  No
 * counterpart in the reference (header-only, compiled into each app). */
const char* gdtp_build_id(void);
int gdtp_device_count(int* count);

/* ---- meshes with per-corner texture coordinates (reference: TexturedMesh,
 *      include/Src/SimpleMesh.h:45-49, read from PLY at :118-160) */
int gdtp_mesh_create(const double* vertices, int num_vertices, const int* triangles, int num_triangles,
                     const double* corner_uv, gdtp_mesh** out);
/* Triangle mesh with per-face texcoord lists from a PLY file (ASCII or binary little-endian); the
 * reference's ReadTexturedMesh (SimpleMesh.h:118-160) without the texture image. */
int gdtp_mesh_read_ply(const char* path, gdtp_mesh** out);
int gdtp_mesh_create_cube(int subdiv, int width, int height, gdtp_mesh** out);
int gdtp_mesh_create_icosphere(int levels, int width, int height, double amplitude, unsigned long long seed, gdtp_mesh** out);
int gdtp_mesh_create_torus(int nu, int nv, int charts_u, int charts_v, double major_radius, double minor_radius,
                           double amplitude, unsigned long long seed, int width, int height, gdtp_mesh** out);
int gdtp_mesh_counts(const gdtp_mesh* mesh, int* num_vertices, int* num_triangles);
int gdtp_mesh_read(const gdtp_mesh* mesh, double* vertices, int* triangles, double* corner_uv);
void gdtp_mesh_destroy(gdtp_mesh* mesh);

/* ---- hierarchy and finest-level mass / stiffness
 * Replaces InitializeHierarchy (HierarchyConstruction.inl:799-849), InitializeBoundaryProlongationData
 * (BoundaryQuadraticElements.inl:775-781), InitializeMetric (Metric.inl:448-466) and
 * InitializeMassAndStiffness (MassAndStiffness.h:490-549).  Block shape 128x16, padding 2x0. */
int gdtp_hierarchy_create(const gdtp_mesh* mesh, int width, int height, int levels, int metric_mode, gdtp_hierarchy** out);
/* The same with a caller-supplied surface metric per mesh triangle: 3 doubles (g00, g01, g11), the
 * metric in the triangle's edge basis (p1 - p0, p2 - p0), used as given (no area normalisation).
 * Replaces InitializeAnisotropicMetric / InitializeVectorFieldMetric + InitializeParameterMetric
 * (Metric.inl:12-67, 415-486) for line integral convolution. */
int gdtp_hierarchy_create_with_metric(const gdtp_mesh* mesh, int width, int height, int levels, const double* triangle_metric,
                                      gdtp_hierarchy** out);
/* Either of the above (triangle_metric NULL: metric_mode) with flags: GDTP_ASSEMBLY_DEVICE assembles
 * the interior cells (every deep row and the interior-cell parts of the boundary rows) on the GPU —
 * the same clipping, quadrature and per-cell triangle order as the host, so the operator agrees
 * with the host assembly to rounding; boundary cells and seams stay on the host.  Cuts mesh
 * admission time.  The reference assembles on the CPU (MassAndStiffness.h:5-549). */
int gdtp_hierarchy_create_ex(const gdtp_mesh* mesh, int width, int height, int levels, int metric_mode, const double* triangle_metric,
                             int flags, gdtp_hierarchy** out);
/* info[0..7] = texels, deep, boundary, charts, raster lines, band tasks, blocks, segments */
int gdtp_hierarchy_level_info(const gdtp_hierarchy* h, int level, int* info);
/* Named integer / double arrays of the hierarchy (see gdtp_capi.cpp for the list); call with
 * out == NULL to get the element count. */
int gdtp_hierarchy_array_i32(const gdtp_hierarchy* h, int level, const char* name, int* out, size_t* count);
int gdtp_hierarchy_array_f64(const gdtp_hierarchy* h, int level, const char* name, double* out, size_t* count);
void gdtp_hierarchy_destroy(gdtp_hierarchy* h);

/* ---- screened-Poisson operator A = screen*M + stiffness*S and its multigrid hierarchy
 * Replaces UpdateLinearSystem (UpdateCoefficients.inl:463-513): Galerkin coarsening, reciprocal
 * deep stencils, exact boundary and coarse LDL^T factorisations.  max_cols = 0 builds the host
 * operator only (no device); otherwise device buffers for up to max_cols right-hand sides. */
int gdtp_system_create(const gdtp_hierarchy* h, double screen, double stiffness, int max_cols, gdtp_system** out);
int gdtp_system_update(gdtp_system* s, double screen, double stiffness);
int gdtp_system_array_i32(const gdtp_system* s, int level, const char* name, int* out, size_t* count);
int gdtp_system_array_f64(const gdtp_system* s, int level, const char* name, double* out, size_t* count);
void gdtp_system_destroy(gdtp_system* s);

/* ---- device operations (FP64, the reference schedule) */
/* VCycle (IterativeSolvers.inl:1211-1333): x (level 0, in/out), b (level 0). */
int gdtp_vcycle(gdtp_system* s, double* x, const double* b, int ncols, void* stream);
/* RelaxationAndResidual (:398-558) when residual != NULL, else Relaxation (:174-395). */
int gdtp_relax(gdtp_system* s, int level, double* x, const double* b, double* residual, int ncols, void* stream);
/* Deep-texel part of one relaxation alone (the "Update Interior" pass of :174-558): the
 * two-coloured band sweep, with the fused deep residual when residual != NULL. */
int gdtp_band_sweep(gdtp_system* s, int level, double* x, const double* b, double* residual, int ncols, void* stream);
/* MultiplyByRestriction (:1009-1071): fine residual of `level` -> coarse rhs of level+1. */
int gdtp_restrict(gdtp_system* s, int level, const double* fine, double* coarse, int ncols, void* stream);
/* AddProlongation (:1122-1168): fine(level) += P coarse(level+1). */
int gdtp_prolong_add(gdtp_system* s, int level, const double* coarse, double* fine, int ncols, void* stream);
/* solve(boundarySolver, ...) (Solver.h:191-203): boundary-indexed vectors of `level`. */
int gdtp_boundary_solve(gdtp_system* s, int level, const double* rhs, double* x, int ncols, void* stream);
/* solve(coarseSolver, ...) (IterativeSolvers.inl:1316). */
int gdtp_coarse_solve(gdtp_system* s, const double* rhs, double* x, int ncols, void* stream);
/* MultiplyBySystemMatrix_NoReciprocals (:881-950) with A, M or S on level 0. */
int gdtp_apply(gdtp_system* s, int which, const double* x, double* y, int ncols, void* stream);
/* r = b - A x on level 0 (ComputeSystemResdiual, :953-1007). */
int gdtp_residual(gdtp_system* s, const double* x, const double* b, double* r, int ncols, void* stream);
/* Per column ||b - A x|| / ||b|| on level 0 in FP64 (host rel_res[ncols], may be NULL) and the two
 * norms (host norms[2 ncols] = ||b_c - A x_c||, ||b_c||, may be NULL), without writing the
 * residual; synchronises `stream`.  Bitwise repeatable (fixed-order reduction).  Replaces
 * ComputeSystemResdiual followed by the apps' residual-norm printout (IterativeSolvers.inl:953-1007,
 * TextureFiltering.cpp:540-558). */
int gdtp_residual_norms(gdtp_system* s, const double* x, const double* b, int ncols, double* rel_res, double* norms, void* stream);
/* Exact FP64 solve of A x = b on level 0: the apps' --useDirectSolver path (FullMatrixConstruction
 * and solve(directSolver, ...), UpdateCoefficients.inl:137-219, TextureFiltering.cpp:509-558,
 * Geodesics.cpp:259-301).  Up to 2^20 texels: LDL^T of the full level-0 matrix (nested dissection,
 * factorised on the host by the first call after create / update and kept; triangular solves on
 * the GPU).  Larger levels: the FP64 direct-equivalent, reference-schedule V-cycles from x = 0
 * until the relative residual reaches 1e-14 or stops falling (the FP64 floor).  x is overwritten;
 * rel_res (host, ncols) and method (host: 0 factorisation, 1 V-cycles) may be NULL. */
int gdtp_direct_solve(gdtp_system* s, double* x, const double* b, int ncols, double* rel_res, int* method, void* stream);
/* gdtp_residual_norms for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): no counterpart in the reference.  Its
 * partial sums live in a static device array of the library (525 KB per GPU, held until the
 * process ends) instead of a buffer allocated on first use, so at most 64 columns
 * (GDTP_ERROR_INVALID beyond).  Calls that overlap on one GPU share that array and race. */
int gdtp_residual_norms_noalloc(gdtp_system* s, const double* x, const double* b, int ncols, double* rel_res, double* norms,
                                void* stream);

/* Normalised-gradient divergence of the heat-method geodesics: for K potentials u (level-0
 * columns), rhs[:, t] = integral < -grad u / |grad u|_g , grad phi_t >_g over the surface, one
 * sample per (mesh triangle ∩ cell) piece (quadratic elements on pieces touching a seam).
 * Replaces IntergrateNormalizedVectorFieldFromSamples (VectorFieldIntergration.inl:452-582) with
 * the boundary prolongation / restriction around it (Geodesics.cpp:522-535).  rhs is overwritten. */
int gdtp_geodesic_divergence(gdtp_system* s, const double* potential, double* rhs, int ncols, void* stream);

/* ---- stiffness modulation (the sharpening brush of TextureFiltering)
 * Cell-based stiffness of fixed signal columns: every active level-0 cell's stiffness applied to
 * each column of `signal` (host, [ncols][N] in global texel order), kept per (texel, cell).
 * Replaces InitializeMassAndStiffness(..., computeCellBasedStiffness = true, ...) and the
 * texelToCellCoeffs / boundaryCellBasedStiffnessRHSMatrix it fills (MassAndStiffness.h:296-395,
 * 490-549; TextureFiltering.cpp:621-649).  Cells are numbered as the hierarchy array "cellInfo". */
int gdtp_modulation_create(const gdtp_hierarchy* h, const double* signal, int ncols, gdtp_modulation** out);
/* out[c][t] = sum over the cells k of texel t of mask[k] * (S_k signal_c)_t, all level-0 texels
 * (mask: device, one weight per cell; out: device [ncols][N], overwritten).  With mask = 1 this is
 * S signal.  Replaces CellStiffnessToTexelStiffness (IterativeSolvers.inl:1170-1208), whose result
 * the filtering rhs adds as M f + alpha * out (TextureFiltering.cpp:569-570). */
int gdtp_modulated_stiffness(const gdtp_modulation* m, const double* mask, double* out, void* stream);
/* info[0..3] = ncols, texels, cells, boundary entries. */
int gdtp_modulation_info(const gdtp_modulation* m, int* info);
/* Device arrays copied to the host (count in / out as for the hierarchy arrays): "deep" (f64,
 * [ncols][4 D]), "boundaryVal" (f64, [ncols][nnz]), "deepCells" (i32, [2 D]), "boundaryRowptr",
 * "boundaryCell" (i32). */
int gdtp_modulation_array_f64(const gdtp_modulation* m, const char* name, double* out, size_t* count);
int gdtp_modulation_array_i32(const gdtp_modulation* m, const char* name, int* out, size_t* count);
void gdtp_modulation_destroy(gdtp_modulation* m);

/* ---- solves to a relative residual (every column: ||b - A x|| <= tol ||b||)
 * mode GDTP_SOLVE_FP64_REFERENCE: reference V-cycles with an FP64 residual check per cycle.
 * mode GDTP_SOLVE_MIXED: FP64 iterative refinement around FP32 V-cycles (four-colour deep
 *   Gauss-Seidel on dense chart planes, restricted additive Schwarz on the boundary rows, exact FP64
 *   coarsest solve),
 *   on the system's own hierarchy; no counterpart in the reference.
 * x holds the initial guess and receives the solution; with GDTP_SOLVE_ZERO_GUESS or-ed into mode
 * the solve starts from x = 0 whatever x holds (it writes the zeros itself, and the mixed path's
 * first residual is b, with no operator applied; no counterpart in the reference).  rel_res (host,
 * ncols) and iterations (host) may be NULL.  Returns GDTP_ERROR_NUMERIC when max_iters is reached
 * first. */
int gdtp_solve(gdtp_system* s, double* x, const double* b, int ncols, double tol, int max_iters, int mode, double* rel_res,
               int* iterations, void* stream);
/* gdtp_solve that calls between(user) on the calling thread after every iteration (mixed: outer
 * iteration; FP64: V-cycle) that leaves the solve unconverged, with the stream idle: a request loop
 * can let other work use the GPU there without restarting the solve (flexible CG keeps its search
 * direction).  between may be NULL.  No counterpart in the reference (its apps step one V-cycle per
 * idle callback, TextureFiltering.cpp:366-376). */
int gdtp_solve_yielding(gdtp_system* s, double* x, const double* b, int ncols, double tol, int max_iters, int mode, double* rel_res,
                        int* iterations, void* stream, void (*between)(void*), void* user);
/* Parametric solve (mixed precision): column c solves (w[2c] M + w[2c+1] S) x = b to `tol` by
 * flexible CG, the FP64 operator applied from the level-0 mass and stiffness with each column's
 * weights and the system's own FP32 V-cycle (built for its create/update weights) as
 * preconditioner; weights is
 * a host array of ncols (screen, stiffness) pairs, screen > 0.  No refactorisation: the per-request
 * operator update of the apps' sliders (UpdateLinearSystem, UpdateCoefficients.inl:463-513) and of
 * batched requests with different weights. */
int gdtp_solve_weighted(gdtp_system* s, double* x, const double* b, int ncols, const double* weights, double tol, int max_iters,
                        double* rel_res, int* iterations, void* stream);
/* Mixed-solver smoothing: pre / post smoothing steps (a pre step: boundary Schwarz, then the four
 * deep colours; the post steps: the deep colours in reverse order, then one boundary Schwarz step
 * after the last), boundary Schwarz sweeps per Schwarz step and the damping of their block update,
 * hipGraph capture of the refinement step (1) or plain launches (0).  Builds the mixed solver's
 * device data if it does not exist yet. */
int gdtp_mixed_configure(gdtp_system* s, int nu_pre, int nu_post, int boundary_sweeps, double omega, int use_graph);
/* Outer iteration of the mixed solver: plain iterative refinement x += V(b - A x) (0, the
 * default: fewest passes over the operator per iteration) or flexible conjugate gradients
 * preconditioned by the FP32 V-cycle (1: fewer iterations on harder operators). */
int gdtp_mixed_acceleration(gdtp_system* s, int flexible_cg);
/* V-cycles per outer iteration of the mixed solver (1..16, or 0 = automatic, the default): with
 * c > 1 the FP32 correction e of A e = r takes c V-cycles (each continuing from the previous e)
 * before the FP64 update and residual, so the FP64 passes over the operator run once per c
 * V-cycles.  Automatic: 2 for plain refinement, 1 for flexible CG and weighted solves and for
 * operators that are nearly singular in FP32 (median row sum / diagonal of level 0 below 2e-6,
 * e.g. the geodesic distance system 1e-6 M + S).  A solve whose
 * residual grows continues with 1.  No counterpart in the reference, whose solve loop is VCycle
 * repeated (TextureFiltering.cpp:562-579). */
int gdtp_mixed_cycles(gdtp_system* s, int cycles);
/* Tile shape of the mixed solver's deep sweep on levels swept by tiles: 0..6 = one fused launch per
 * sweep (all four colours, the tile and its halo staged in LDS; the shapes' tile sizes, threads
 * and variants are listed by gdtp_mixed_tile_shape_info); -1 = one launch per colour on 64x4
 * tiles; GDTP_TILE_SHAPE_PAIRS = one launch per colour pair ((0, 1) then (2, 3); post-smoothing
 * (3, 2) then (1, 0)), a workgroup per row of the colour planes (levels whose widest chart row
 * exceeds 2048 sub-grid texels fall back to one launch per colour);
 * GDTP_TILE_SHAPE_AUTO (the default) = per level, one launch per colour where a colour plane
 * holds more than 2^21 texels, colour pairs above 2^17, else shape 2.  Every shape gives bitwise the same iterate.  Setting a
 * shape here replaces any per-level shapes.  shape -2 only reports; `current` (may be NULL)
 * receives the setting (GDTP_TILE_SHAPE_AUTO included).  No counterpart in the reference. */
int gdtp_mixed_tile_shape(gdtp_system* s, int shape, int* current);
/* Fused tile shape `shape` (0..6): tile width x height in sub-grid texels (one texel of each of
 * the four colour planes), threads per workgroup, and the variant bits: 3 = prefetch depth (0:
 * each colour's operands fetched right before it, 1: one colour ahead, 2: all four colours at the
 * start), 4 = XCD-contiguous tile order.  Any pointer may be NULL. */
int gdtp_mixed_tile_shape_info(int shape, int* width, int* height, int* threads, int* variant);
/* The same for one level only (the best shape differs between levels); shape -2 only reports;
 * `current` receives the shape the level uses (-4, -1..6, the automatic choice resolved). */
int gdtp_mixed_level_tile_shape(gdtp_system* s, int level, int shape, int* current);
/* Transfers fused into the mixed solver's relaxation kernels: GDTP_FUSE_PROLONGATION lets the first
 * post-smoothing sweep of every level swept by fused tiles add the bilinear prolongation of the
 * coarse correction itself, instead of a separate pass over the level; 0 (the default) keeps the
 * separate pass, which measured faster on every level at C2 (level 1: 77 against 87 us for the
 * prolongation and the first post sweep).  Both give bitwise the same iterate.  No counterpart in
 * the reference (AddProlongation then Relaxation, IterativeSolvers.inl:1122-1168, 174-395). */
int gdtp_mixed_fuse_transfers(gdtp_system* s, int flags);
/* One four-colour deep Gauss-Seidel sweep of the mixed solver on `level` (its internal FP32
 * vectors); exposed to time the dominant kernel. */
int gdtp_mixed_deep_sweep(gdtp_system* s, int level, int ncols, void* stream);
/* Device bytes held by the system (reference path, mixed path). */
int gdtp_system_memory(const gdtp_system* s, size_t* reference_bytes, size_t* mixed_bytes);

#ifdef __cplusplus
}
#endif
#endif /* GDTP_H */
