// C ABI of libgdtp (declared in include/gdtp.h).
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/gdtp.h"
#include "gdtp_device.h"
#include "gdtp_mixed.h"
#include "gdtp_mesh.h"

using namespace gdtp;

struct gdtp_mesh {
  TexturedMesh m;
};

struct gdtp_hierarchy {
  TexturedMesh mesh;
  Hierarchy h;
  FineOperator op;
  std::vector<std::vector<M2>> metric;                 // per chart triangle, UV coordinates
  mutable std::unique_ptr<VectorFieldSamples> vf;      // built on first use
  mutable std::mutex vfMutex;
};

static const VectorFieldSamples& Samples(const gdtp_hierarchy* h) {
  std::lock_guard<std::mutex> lock(h->vfMutex);
  if (!h->vf) {
    auto vs = std::make_unique<VectorFieldSamples>();
    std::string err;
    if (!BuildVectorFieldSamples(h->h, h->metric, *vs, err)) throw std::runtime_error(err);
    h->vf = std::move(vs);
  }
  return *h->vf;
}

struct gdtp_modulation {
  DevModulation d;       // on the GPU when one is present
  bool onDevice = false;
  CellStiffnessRhs host;  // kept on the host only without a GPU (set-up checks on CPU machines)
  int N = 0, D = 0, B = 0;
};

// LDL^T factor of the full level-0 matrix for gdtp_direct_solve (host factor, GPU triangular solves).
struct DirectFactor {
  Ldlt F;
  DevLdlt d;
  double* y = nullptr;  // [kcap][N] solve scratch
  std::vector<void*> allocations;
  ~DirectFactor() {
    for (void* p : allocations) (void)hipFree(p);
  }
};

struct gdtp_system {
  const gdtp_hierarchy* H = nullptr;
  double screen = 1, stiff = 0;
  int kcap = 0;
  OperatorHierarchy oh;
  std::unique_ptr<DeviceSystem> ds;
  std::unique_ptr<MixedSystem> mixed;
  std::unique_ptr<DevVectorField> vf;
  std::unique_ptr<DirectFactor> direct;  // built by the first gdtp_direct_solve after create / update
  int nuPre = 1, nuPost = 2, boundarySweeps = 1, useGraph = 1, krylov = 0, cycles = 0, tileShape = GDTP_TILE_SHAPE_AUTO, fuseTransfers = 0;
  double omega = 1.0;
  std::vector<int> levelShape;  // per-level deep-sweep tile shapes set since the last all-level setting
  double* dNorms = nullptr;
  double* hNorms = nullptr;
  double* dNormParts = nullptr;  // residual-norm partial sums, NormPartsSize(kcap) + 2 kcap doubles
  ~gdtp_system() {
    if (dNorms) (void)hipFree(dNorms);
    if (dNormParts) (void)hipFree(dNormParts);
    if (hNorms) (void)hipHostFree(hNorms);
  }
};

static thread_local std::string g_error;

static int Fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define GDTP_TRY(body)                                                        \
  try {                                                                       \
    body                                                                      \
  } catch (const std::bad_alloc&) {                                           \
    return Fail(GDTP_ERROR_DEVICE, "out of memory");                          \
  } catch (const std::exception& e) {                                         \
    return Fail(GDTP_ERROR_DEVICE, e.what());                                 \
  }

// Definitions take C linkage from the declarations in gdtp.h.

const char* gdtp_last_error(void) { return g_error.c_str(); }

#ifndef GDTP_BUILD_ID
#define GDTP_BUILD_ID "unknown"
#endif
// kept whole in the library's read-only data, so that the identifier can be read from the file
// without loading it
static const char kBuildTag[] = "gdtp-build-id:" GDTP_BUILD_ID;
const char* gdtp_build_id(void) { return kBuildTag + 14; }

int gdtp_device_count(int* count) {
  if (!count) return Fail(GDTP_ERROR_INVALID, "null count");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return GDTP_OK;
}

// ---------------------------------------------------------------- meshes
int gdtp_mesh_create(const double* V, int nv, const int* T, int nt, const double* uv, gdtp_mesh** out) {
  if (!V || !T || !uv || !out || nv <= 0 || nt <= 0) return Fail(GDTP_ERROR_INVALID, "invalid mesh arrays");
  GDTP_TRY({
    auto* m = new gdtp_mesh();
    m->m.vertices.resize(nv);
    for (int i = 0; i < nv; i++) m->m.vertices[i] = V3(V[3 * i], V[3 * i + 1], V[3 * i + 2]);
    m->m.triangles.assign(T, T + 3 * (size_t)nt);
    for (size_t k = 0; k < 3 * (size_t)nt; k++)
      if (T[k] < 0 || T[k] >= nv) { delete m; return Fail(GDTP_ERROR_INVALID, "triangle index out of range"); }
    m->m.uv.resize(3 * (size_t)nt);
    for (size_t k = 0; k < 3 * (size_t)nt; k++) m->m.uv[k] = V2(uv[2 * k], uv[2 * k + 1]);
    *out = m;
    return GDTP_OK;
  })
}

int gdtp_mesh_read_ply(const char* path, gdtp_mesh** out) {
  if (!path || !out) return Fail(GDTP_ERROR_INVALID, "null argument");
  GDTP_TRY({
    auto* m = new gdtp_mesh();
    std::string err;
    if (!ReadPly(path, m->m, err)) { delete m; return Fail(GDTP_ERROR_GEOMETRY, err); }
    *out = m;
    return GDTP_OK;
  })
}

int gdtp_mesh_create_cube(int subdiv, int W, int H, gdtp_mesh** out) {
  if (!out) return Fail(GDTP_ERROR_INVALID, "null output");
  auto* m = new gdtp_mesh();
  if (!MakeCube(subdiv, W, H, m->m)) { delete m; return Fail(GDTP_ERROR_INVALID, "invalid cube parameters"); }
  *out = m;
  return GDTP_OK;
}

int gdtp_mesh_create_icosphere(int levels, int W, int H, double amplitude, unsigned long long seed, gdtp_mesh** out) {
  if (!out) return Fail(GDTP_ERROR_INVALID, "null output");
  auto* m = new gdtp_mesh();
  if (!MakeIcosphere(levels, W, H, amplitude, seed, m->m)) { delete m; return Fail(GDTP_ERROR_INVALID, "invalid icosphere parameters"); }
  *out = m;
  return GDTP_OK;
}

int gdtp_mesh_create_torus(int nu, int nv, int cu, int cv, double R, double r, double amplitude, unsigned long long seed, int W, int H,
                           gdtp_mesh** out) {
  if (!out) return Fail(GDTP_ERROR_INVALID, "null output");
  auto* m = new gdtp_mesh();
  if (!MakeTorus(nu, nv, cu, cv, R, r, amplitude, seed, W, H, m->m)) { delete m; return Fail(GDTP_ERROR_INVALID, "invalid torus parameters"); }
  *out = m;
  return GDTP_OK;
}

int gdtp_mesh_counts(const gdtp_mesh* m, int* nv, int* nt) {
  if (!m) return Fail(GDTP_ERROR_INVALID, "null mesh");
  if (nv) *nv = (int)m->m.vertices.size();
  if (nt) *nt = m->m.numTriangles();
  return GDTP_OK;
}

int gdtp_mesh_read(const gdtp_mesh* m, double* V, int* T, double* uv) {
  if (!m) return Fail(GDTP_ERROR_INVALID, "null mesh");
  if (V) memcpy(V, m->m.vertices.data(), m->m.vertices.size() * sizeof(V3));
  if (T) memcpy(T, m->m.triangles.data(), m->m.triangles.size() * sizeof(int));
  if (uv) memcpy(uv, m->m.uv.data(), m->m.uv.size() * sizeof(V2));
  return GDTP_OK;
}

void gdtp_mesh_destroy(gdtp_mesh* m) { delete m; }

// ---------------------------------------------------------------- hierarchy
// Shared by the hierarchy constructors: metric from the embedding / UV layout, or the given
// per-triangle surface metric; interior cells assembled on the GPU with `device`.
static int CreateHierarchy(const gdtp_mesh* mesh, int W, int H, int levels, int metricMode, const double* triangleMetric, bool device,
                           gdtp_hierarchy** out) {
  if (!mesh || !out || W < 8 || H < 8 || levels < 1 || levels > 16) return Fail(GDTP_ERROR_INVALID, "invalid hierarchy arguments");
  std::vector<M2> g;
  if (triangleMetric) {
    const int nT = mesh->m.numTriangles();
    g.resize(nT);
    for (int t = 0; t < nT; t++) {
      g[t](0, 0) = triangleMetric[3 * t], g[t](0, 1) = g[t](1, 0) = triangleMetric[3 * t + 1], g[t](1, 1) = triangleMetric[3 * t + 2];
      if (!(g[t](0, 0) > 0) || !(g[t].det() > 0)) return Fail(GDTP_ERROR_INVALID, "triangle metric is not positive definite");
    }
  }
  GDTP_TRY({
    auto* h = new gdtp_hierarchy();
    h->mesh = mesh->m;
    std::string err;
    BlockInfo bi;
    if (!BuildHierarchy(h->mesh, W, H, levels, bi, h->h, err)) { delete h; return Fail(GDTP_ERROR_GEOMETRY, err); }
    if (triangleMetric) ComputeParameterMetricFromSurface(h->mesh, h->h.atlas, g, h->metric);
    else ComputeParameterMetric(h->mesh, h->h.atlas, metricMode == GDTP_METRIC_UNIFORM ? METRIC_UNIFORM : METRIC_EMBEDDING, h->metric);
    if (!AssembleFineOperator(h->mesh, h->h, h->metric, h->op, err, nullptr, 0, nullptr, device)) {
      delete h;
      return Fail(GDTP_ERROR_GEOMETRY, err);
    }
    *out = h;
    return GDTP_OK;
  })
}

int gdtp_hierarchy_create(const gdtp_mesh* mesh, int W, int H, int levels, int metricMode, gdtp_hierarchy** out) {
  return CreateHierarchy(mesh, W, H, levels, metricMode, nullptr, false, out);
}

int gdtp_hierarchy_create_with_metric(const gdtp_mesh* mesh, int W, int H, int levels, const double* triangleMetric, gdtp_hierarchy** out) {
  if (!triangleMetric) return Fail(GDTP_ERROR_INVALID, "null triangle metric");
  return CreateHierarchy(mesh, W, H, levels, GDTP_METRIC_EMBEDDING, triangleMetric, false, out);
}

int gdtp_hierarchy_create_ex(const gdtp_mesh* mesh, int W, int H, int levels, int metricMode, const double* triangleMetric, int flags,
                             gdtp_hierarchy** out) {
  if (flags & ~GDTP_ASSEMBLY_DEVICE) return Fail(GDTP_ERROR_INVALID, "unknown hierarchy flags");
  return CreateHierarchy(mesh, W, H, levels, metricMode, triangleMetric, (flags & GDTP_ASSEMBLY_DEVICE) != 0, out);
}

int gdtp_hierarchy_level_info(const gdtp_hierarchy* h, int level, int* info) {
  if (!h || !info || level < 0 || level >= (int)h->h.levels.size()) return Fail(GDTP_ERROR_INVALID, "invalid level");
  const Level& L = h->h.levels[level];
  info[0] = L.numTexels, info[1] = L.numDeep, info[2] = L.numBoundary, info[3] = (int)L.charts.size();
  info[4] = (int)L.rasterLines.size(), info[5] = (int)L.tasks.taskPtr.size() - 1, info[6] = (int)L.tasks.blockBegin.size();
  info[7] = (int)L.tasks.segments.size();
  return GDTP_OK;
}

template <class T>
static int Emit(const std::vector<T>& v, T* out, size_t* count) {
  if (!count) return Fail(GDTP_ERROR_INVALID, "null count");
  if (out) {
    if (*count < v.size()) return Fail(GDTP_ERROR_INVALID, "output buffer too small");
    if (!v.empty()) memcpy(out, v.data(), v.size() * sizeof(T));
  }
  *count = v.size();
  return GDTP_OK;
}

static bool CsrI32(const Csr& m, const std::string& name, const std::string& base, std::vector<int>& v) {
  if (name == base + ".rowptr") return v = m.rowptr, true;
  if (name == base + ".col") return v = m.col, true;
  return false;
}

int gdtp_hierarchy_array_i32(const gdtp_hierarchy* h, int level, const char* cname, int* out, size_t* count) {
  if (!h || !cname || level < 0 || level >= (int)h->h.levels.size()) return Fail(GDTP_ERROR_INVALID, "invalid arguments");
  if (std::strncmp(cname, "vf.", 3) == 0 && level == 0) {  // samples are built on first use
    try {
      Samples(h);
    } catch (const std::exception& e) {
      return Fail(GDTP_ERROR_GEOMETRY, e.what());
    }
  }
  const Level& L = h->h.levels[level];
  const std::string name(cname);
  std::vector<int> v;
  if (name == "boundaryGlobalIndex") v = L.boundaryGlobalIndex;
  else if (name == "deepGlobalIndex") v = L.deepGlobalIndex;
  else if (name == "boundaryAndDeepIndex") v = L.boundaryAndDeepIndex;
  else if (name == "nodeInfo") { for (auto& n : L.nodeInfo) v.insert(v.end(), {n.chart, n.ci, n.cj, n.type}); }
  else if (name == "rasterLines") { for (auto& x : L.rasterLines) v.insert(v.end(), {x.start, x.end, x.prev, x.next, x.coeff}); }
  else if (name == "restrictionLines") { for (auto& x : L.restrictionLines) v.insert(v.end(), {x.start, x.end, x.prev, x.next, x.coeff}); }
  else if (name == "deepLines") { for (auto& x : L.deepLines) v.insert(v.end(), {x.coarseStart, x.coarseEnd, x.finePrev, x.fineCurrent, x.fineNext}); }
  else if (name == "prolongationLines") { for (auto& x : L.prolongationLines) v.insert(v.end(), {x.start, x.length, x.center, x.prev, x.next, x.aligned}); }
  else if (name == "tasks.taskPtr") v = L.tasks.taskPtr;
  else if (name == "tasks.blockBegin") v = L.tasks.blockBegin;
  else if (name == "tasks.blockDeepEnd") v = L.tasks.blockDeepEnd;
  else if (name == "tasks.blockPadEnd") v = L.tasks.blockPadEnd;
  else if (name == "tasks.linePtr") v = L.tasks.linePtr;
  else if (name == "tasks.lineRow") v = L.tasks.lineRow;
  else if (name == "tasks.blockStart") v = L.tasks.blockStart;
  else if (name == "tasks.taskSteps") v = L.tasks.taskSteps;
  else if (name == "tasks.lineElem") v = L.tasks.lineElem;
  else if (name == "tasks.elems") v = L.tasks.elems;
  else if (name == "tasks.segments") { for (auto& s : L.tasks.segments) v.insert(v.end(), {s.currentStart, s.currentEnd, s.previousStart, s.nextStart, s.deepStart}); }
  else if (name == "charts") {
    for (auto& c : L.charts)
      v.insert(v.end(), {c.width, c.height, c.cornerCoords[0], c.cornerCoords[1], c.centerOffset[0], c.centerOffset[1], c.texelOffset, c.deepOffset, c.boundaryOffset});
  } else if (name == "nodeType") { for (auto& c : L.charts) for (auto t : c.nodeType) v.push_back(t); }
  else if (name == "cellType") { for (auto& c : L.charts) for (auto t : c.cellType) v.push_back(t); }
  else if (name == "chartTexelIndex") { for (auto& c : L.charts) v.insert(v.end(), c.globalTexel.begin(), c.globalTexel.end()); }
  else if (name == "cellInfo") {  // per active cell in global order: chart, i, j, type (0 boundary, 1 interior)
    v.assign(4 * (size_t)L.numCells, 0);
    for (int ch = 0; ch < (int)L.charts.size(); ch++) {
      const GridChart& g = L.charts[ch];
      for (int j = 0; j + 1 < g.height; j++)
        for (int i = 0; i + 1 < g.width; i++) {
          const int lc = g.localCell[(size_t)j * (g.width - 1) + i];
          if (lc < 0) continue;
          const size_t at = 4 * (size_t)(g.cellOffset + lc);
          v[at] = ch, v[at + 1] = i, v[at + 2] = j, v[at + 3] = g.cell(i, j);
        }
    }
  }
  else if (CsrI32(L.boundaryRestriction, name, "boundaryRestriction", v)) {
  } else if (level == 0 && name == "vf.cellCorners") v = Samples(h).cellCorners;
  else if (level == 0 && name == "vf.genPtr") v = Samples(h).genPtr;
  else if (level == 0 && name == "vf.genTexel") v = Samples(h).genTexel;
  else if (level == 0 && (CsrI32(h->op.bbMass, name, "mass.bb", v) || CsrI32(h->op.bbStiffness, name, "stiffness.bb", v) ||
                            CsrI32(h->op.bdMass, name, "mass.bd", v) || CsrI32(h->op.bdStiffness, name, "stiffness.bd", v))) {
  } else return Fail(GDTP_ERROR_INVALID, "unknown hierarchy array " + name);
  return Emit(v, out, count);
}

int gdtp_hierarchy_array_f64(const gdtp_hierarchy* h, int level, const char* cname, double* out, size_t* count) {
  if (!h || !cname || level < 0 || level >= (int)h->h.levels.size()) return Fail(GDTP_ERROR_INVALID, "invalid arguments");
  if (std::strncmp(cname, "vf.", 3) == 0 && level == 0) {  // samples are built on first use
    try {
      Samples(h);
    } catch (const std::exception& e) {
      return Fail(GDTP_ERROR_GEOMETRY, e.what());
    }
  }
  const Level& L = h->h.levels[level];
  const std::string name(cname);
  std::vector<double> v;
  if (name == "boundaryRestriction.val") v = L.boundaryRestriction.val;
  else if (name == "chartsGeom") { for (auto& c : L.charts) v.insert(v.end(), {c.corner.x, c.corner.y, c.cellW, c.cellH}); }
  else if (level == 0 && name.rfind("vf.", 0) == 0) {
    const VectorFieldSamples& vs = Samples(h);
    const std::vector<float>* f = name == "vf.cellData" ? &vs.cellData : name == "vf.genTensor" ? &vs.genTensor
                                : name == "vf.genG" ? &vs.genG : name == "vf.genV" ? &vs.genV : nullptr;
    if (f) v.assign(f->begin(), f->end());
    else if (name == "vf.invCell") v = {vs.invCellW, vs.invCellH};
    else return Fail(GDTP_ERROR_INVALID, "unknown hierarchy array " + name);
  }
  else if (level == 0 && name == "mass.deep") v = h->op.deepMass;
  else if (level == 0 && name == "stiffness.deep") v = h->op.deepStiffness;
  else if (level == 0 && name == "mass.bb.val") v = h->op.bbMass.val;
  else if (level == 0 && name == "stiffness.bb.val") v = h->op.bbStiffness.val;
  else if (level == 0 && name == "mass.bd.val") v = h->op.bdMass.val;
  else if (level == 0 && name == "stiffness.bd.val") v = h->op.bdStiffness.val;
  else return Fail(GDTP_ERROR_INVALID, "unknown hierarchy array " + name);
  return Emit(v, out, count);
}

void gdtp_hierarchy_destroy(gdtp_hierarchy* h) { delete h; }

// ---------------------------------------------------------------- system
static int BuildSystem(gdtp_system* s, double screen, double stiff) {
  std::string err;
  OperatorHierarchy oh;
  if (!BuildOperatorHierarchy(s->H->h, s->H->op, screen, stiff, true, oh, err)) return Fail(GDTP_ERROR_NUMERIC, err);
  std::unique_ptr<DeviceSystem> ds;
  if (s->kcap > 0) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return Fail(GDTP_ERROR_DEVICE, "no HIP device available");
    ds.reset(new DeviceSystem());
    UploadSystem(s->H->h, oh, s->H->op, s->kcap, *ds);
  }
  s->mixed.reset();
  s->direct.reset();
  s->oh = std::move(oh);
  s->ds = std::move(ds);
  s->screen = screen, s->stiff = stiff;
  return GDTP_OK;
}

int gdtp_system_create(const gdtp_hierarchy* h, double screen, double stiff, int maxCols, gdtp_system** out) {
  if (!h || !out || maxCols < 0 || maxCols > 4096) return Fail(GDTP_ERROR_INVALID, "invalid system arguments");
  GDTP_TRY({
    auto* s = new gdtp_system();
    s->H = h;
    s->kcap = maxCols;
    int rc = BuildSystem(s, screen, stiff);
    if (rc != GDTP_OK) { delete s; return rc; }
    *out = s;
    return GDTP_OK;
  })
}

int gdtp_system_update(gdtp_system* s, double screen, double stiff) {
  if (!s) return Fail(GDTP_ERROR_INVALID, "null system");
  GDTP_TRY({ return BuildSystem(s, screen, stiff); })
}

int gdtp_system_array_i32(const gdtp_system* s, int level, const char* cname, int* out, size_t* count) {
  if (!s || !cname || level < 0 || level >= (int)s->oh.levels.size()) return Fail(GDTP_ERROR_INVALID, "invalid arguments");
  const std::string name(cname);
  const LevelOperator& O = s->oh.levels[level];
  std::vector<int> v;
  if (CsrI32(O.bd, name, "bd", v) || CsrI32(O.bb, name, "bb", v) || CsrI32(s->oh.coarseMatrix, name, "coarse", v)) {
  } else if (name == "boundaryFactor.perm" && level + 1 < (int)s->oh.levels.size()) v = s->oh.boundaryFactor[level].perm;
  else if (name == "coarseFactor.perm") v = s->oh.coarseFactor.perm;
  else return Fail(GDTP_ERROR_INVALID, "unknown system array " + name);
  return Emit(v, out, count);
}

int gdtp_system_array_f64(const gdtp_system* s, int level, const char* cname, double* out, size_t* count) {
  if (!s || !cname || level < 0 || level >= (int)s->oh.levels.size()) return Fail(GDTP_ERROR_INVALID, "invalid arguments");
  const std::string name(cname);
  const LevelOperator& O = s->oh.levels[level];
  std::vector<double> v;
  if (name == "deepRaw") v = O.deepRaw;
  else if (name == "deepRecip") v = O.deepRecip;
  else if (name == "bd.val") v = O.bd.val;
  else if (name == "bb.val") v = O.bb.val;
  else if (name == "coarse.val") v = s->oh.coarseMatrix.val;
  else return Fail(GDTP_ERROR_INVALID, "unknown system array " + name);
  return Emit(v, out, count);
}

void gdtp_system_destroy(gdtp_system* s) { delete s; }

// ---------------------------------------------------------------- device operations
static int Ready(gdtp_system* s, int ncols) {
  if (!s) return Fail(GDTP_ERROR_INVALID, "null system");
  if (!s->ds) return Fail(GDTP_ERROR_INVALID, "system was created without device buffers (max_cols = 0)");
  if (ncols < 1 || ncols > s->kcap) return Fail(GDTP_ERROR_INVALID, "ncols outside [1, max_cols]");
  return GDTP_OK;
}

#define GDTP_DEVICE_OP(cond, body)                                      \
  int rc_ = Ready(s, ncols);                                            \
  if (rc_ != GDTP_OK) return rc_;                                       \
  if (!(cond)) return Fail(GDTP_ERROR_INVALID, "invalid arguments");    \
  GDTP_TRY({                                                            \
    hipStream_t st = (hipStream_t)stream;                               \
    DeviceSystem& ds = *s->ds;                                          \
    body;                                                               \
    GDTP_HIP(hipGetLastError());                                        \
    return GDTP_OK;                                                     \
  })

int gdtp_vcycle(gdtp_system* s, double* x, const double* b, int ncols, void* stream) {
  GDTP_DEVICE_OP(x && b, VCycleReference(ds, x, b, ncols, ds.L[0].N, st))
}

int gdtp_relax(gdtp_system* s, int level, double* x, const double* b, double* residual, int ncols, void* stream) {
  GDTP_DEVICE_OP(x && b && level >= 0 && level + 1 < s->ds->levels,
                 RelaxLevel(ds, level, x, b, residual, ncols, ds.L[level].N, residual != nullptr, st))
}

int gdtp_band_sweep(gdtp_system* s, int level, double* x, const double* b, double* residual, int ncols, void* stream) {
  GDTP_DEVICE_OP(x && b && level >= 0 && level + 1 < s->ds->levels,
                 BandSweepLevel(ds, level, x, b, residual, ncols, ds.L[level].N, st))
}

int gdtp_restrict(gdtp_system* s, int level, const double* fine, double* coarse, int ncols, void* stream) {
  GDTP_DEVICE_OP(fine && coarse && level >= 0 && level + 1 < s->ds->levels,
                 RestrictLevel(ds, level, fine, coarse, ncols, ds.L[level].N, ds.L[level + 1].N, st))
}

int gdtp_prolong_add(gdtp_system* s, int level, const double* coarse, double* fine, int ncols, void* stream) {
  GDTP_DEVICE_OP(fine && coarse && level >= 0 && level + 1 < s->ds->levels,
                 ProlongAddLevel(ds, level, coarse, fine, ncols, ds.L[level + 1].N, ds.L[level].N, st))
}

int gdtp_boundary_solve(gdtp_system* s, int level, const double* rhs, double* x, int ncols, void* stream) {
  GDTP_DEVICE_OP(rhs && x && level >= 0 && level + 1 < s->ds->levels,
                 BoundarySolveLevel(ds, level, rhs, x, ncols, ds.L[level].B, st))
}

int gdtp_coarse_solve(gdtp_system* s, const double* rhs, double* x, int ncols, void* stream) {
  GDTP_DEVICE_OP(rhs && x, CoarseSolve(ds, rhs, x, ncols, ds.L[ds.levels - 1].N, st))
}

int gdtp_apply(gdtp_system* s, int which, const double* x, double* y, int ncols, void* stream) {
  GDTP_DEVICE_OP(x && y && which >= 0 && which <= 2,
                 ApplyOperator(ds, which == 0 ? ds.L[0] : (which == 1 ? ds.mass0 : ds.stiff0), x, y, ncols, ds.L[0].N, st))
}

int gdtp_residual(gdtp_system* s, const double* x, const double* b, double* r, int ncols, void* stream) {
  GDTP_DEVICE_OP(x && b && r, ResidualLevel(ds, 0, x, b, r, ncols, ds.L[0].N, st))
}

// Host side of the residual norms: squared sums (device, 2 K doubles) -> norms and relative residuals.
static void FinishNorms(const double* sq, int K, double* relRes, double* norms) {
  for (int c = 0; c < K; c++) {
    const double r = std::sqrt(sq[2 * c]), bn = std::sqrt(sq[2 * c + 1]);
    if (relRes) relRes[c] = r / std::max(bn, 1e-300);
    if (norms) norms[2 * c] = r, norms[2 * c + 1] = bn;
  }
}

// Squared residual / rhs norms of K level-0 columns into the host vector sq (2 K entries), through
// the system's norm scratch (allocated on first use).
static void SystemNorms(gdtp_system* s, const double* x, const double* b, int K, std::vector<double>& sq, hipStream_t st) {
  DeviceSystem& ds = *s->ds;
  if (!s->dNormParts) GDTP_HIP(hipMalloc(&s->dNormParts, (NormPartsSize(s->kcap) + 2 * (size_t)s->kcap) * sizeof(double)));
  double* out = s->dNormParts + NormPartsSize(s->kcap);
  ResidualNorms(ds, x, b, K, ds.L[0].N, s->dNormParts, out, st);
  sq.assign(2 * (size_t)K, 0.0);
  GDTP_HIP(hipMemcpyAsync(sq.data(), out, sq.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  GDTP_HIP(hipStreamSynchronize(st));
}

int gdtp_residual_norms(gdtp_system* s, const double* x, const double* b, int ncols, double* rel_res, double* norms, void* stream) {
  GDTP_DEVICE_OP(x && b, {
    std::vector<double> sq;
    SystemNorms(s, x, b, ncols, sq, st);
    FinishNorms(sq.data(), ncols, rel_res, norms);
  })
}

// Largest level 0 that gdtp_direct_solve factorises (the host factorisation grows as N^1.5).
static constexpr int kDirectMaxTexels = 1 << 20;

int gdtp_direct_solve(gdtp_system* s, double* x, const double* b, int ncols, double* rel_res, int* method, void* stream) {
  GDTP_DEVICE_OP(x && b, {
    const int N = ds.L[0].N;
    std::vector<double> sq;
    if (N <= kDirectMaxTexels) {
      if (!s->direct) {
        auto f = std::make_unique<DirectFactor>();
        const LevelOperator& op = s->oh.levels[0];
        const Csr A = FullLevelMatrix(s->H->h.levels[0], op.deepRaw, op.bb, op.bd);
        std::string err;
        if (!FactorLdlt(A, f->F, err)) return Fail(GDTP_ERROR_NUMERIC, "direct factorisation: " + err);
        UploadLdlt(f->allocations, f->F, f->d, s->kcap);
        GDTP_HIP(hipMalloc(&f->y, (size_t)s->kcap * N * sizeof(double)));
        f->allocations.push_back(f->y);
        s->direct = std::move(f);
      }
      LdltSolve(s->direct->d, b, N, x, N, nullptr, s->direct->y, ncols, st);
      SystemNorms(s, x, b, ncols, sq, st);
      if (method) *method = 0;
    } else {
      // FP64 direct-equivalent: reference V-cycles from zero until the worst column's relative
      // residual reaches 1e-14 or has not fallen below 0.9 of its best for three cycles
      GDTP_HIP(hipMemsetAsync(x, 0, (size_t)ncols * N * sizeof(double), st));
      double best = 1e300;
      for (int it = 0, stall = 0; it < 300 && stall < 3; it++) {
        VCycleReference(ds, x, b, ncols, N, st);
        SystemNorms(s, x, b, ncols, sq, st);
        double worst = 0.0;
        for (int c = 0; c < ncols; c++) worst = std::max(worst, std::sqrt(sq[2 * c]) / std::max(std::sqrt(sq[2 * c + 1]), 1e-300));
        if (worst <= 1e-14) break;
        if (worst < 0.9 * best) {
          best = worst;
          stall = 0;
        } else {
          stall++;
        }
      }
      if (method) *method = 1;
    }
    FinishNorms(sq.data(), ncols, rel_res, nullptr);
  })
}

int gdtp_residual_norms_noalloc(gdtp_system* s, const double* x, const double* b, int ncols, double* rel_res, double* norms,
                                void* stream) {
  if (s && s->ds && ncols > kNoAllocMaxCols)
    return Fail(GDTP_ERROR_INVALID, "ncols exceeds the 64 columns of the no-allocation scratch");
  GDTP_DEVICE_OP(x && b, {
    double sq[2 * kNoAllocMaxCols];
    ResidualNormsNoAlloc(ds, x, b, ncols, ds.L[0].N, sq, st);
    FinishNorms(sq, ncols, rel_res, norms);
  })
}

// The automatic deep-sweep tile shape of a level: one launch per colour where a colour plane holds
// more than 2^21 sub-grid texels (there the per-colour sweep streams at the memory's rate and the
// fused one does not keep up: C2 level 0, 3.7 M texels per plane, 186 against 188-192 us per sweep),
// else the fused 64x4 tiles of 384 threads, one colour ahead, XCD-contiguous (C2 levels 1-3: 46.7,
// 17.6 and 8.3 us per sweep against 58.0, 19.7 and 14.6 us per colour).
// Measured at C2 (us per sweep; texels per dense colour plane): level 0 (6.2 M) one launch per
// colour 187, fused 190, pairs 211; level 1 (1.6 M) pairs 42, fused shape 2 48; level 2 (0.41 M)
// pairs 16, shape 2 18; level 3 (0.12 M) fused 8, pairs 10.
static int AutoTileShape(const DevMixLevel& d) {
  return d.plane > (1L << 21) ? kShapePerColour : d.plane > (1L << 17) ? kShapePairs : 2;
}

static MixedSystem& Mixed(gdtp_system* s) {
  if (!s->mixed) {
    std::unique_ptr<MixedSystem> m(new MixedSystem());
    UploadMixed(s->H->h, s->oh, *s->ds, s->kcap, *m);
    s->mixed = std::move(m);
  }
  MixedSystem& m = *s->mixed;
  m.nuPre = s->nuPre, m.nuPost = s->nuPost, m.boundarySweeps = s->boundarySweeps, m.omega = (float)s->omega;
  m.screen = s->screen, m.stiff = s->stiff, m.krylov = s->krylov != 0, m.cycles = s->cycles;
  for (int l = 0; l < m.levels; l++) {
    const int sh = l < (int)s->levelShape.size() ? s->levelShape[l] : s->tileShape;
    m.L[l].shape = sh == GDTP_TILE_SHAPE_AUTO ? AutoTileShape(m.L[l]) : sh;
  }
  m.fuseProlong = (s->fuseTransfers & GDTP_FUSE_PROLONGATION) != 0;
  return m;
}

int gdtp_mixed_fuse_transfers(gdtp_system* s, int flags) {
  if (!s || (flags & ~GDTP_FUSE_PROLONGATION)) return Fail(GDTP_ERROR_INVALID, "unknown transfer fusion flags");
  s->fuseTransfers = flags;
  if (s->mixed) {
    s->mixed->fuseProlong = (flags & GDTP_FUSE_PROLONGATION) != 0;
    InvalidateGraph(*s->mixed);
  }
  return GDTP_OK;
}

// A deep-sweep shape setting: a fused shape, one launch per colour, colour pairs, automatic, or
// (report only) -2.
static bool ValidShape(int shape) {
  return (shape >= -2 && shape < kFusedShapes) || shape == GDTP_TILE_SHAPE_AUTO || shape == GDTP_TILE_SHAPE_PAIRS;
}

int gdtp_mixed_tile_shape(gdtp_system* s, int shape, int* current) {
  if (!s || !ValidShape(shape)) return Fail(GDTP_ERROR_INVALID, "tile shape outside [-4, 7)");
  if (shape != -2) {
    s->tileShape = shape;
    s->levelShape.clear();
    if (s->mixed) {
      for (DevMixLevel& d : s->mixed->L) d.shape = shape == GDTP_TILE_SHAPE_AUTO ? AutoTileShape(d) : shape;
      InvalidateGraph(*s->mixed);
    }
  }
  if (current) *current = s->tileShape;
  return GDTP_OK;
}

int gdtp_mixed_tile_shape_info(int shape, int* width, int* height, int* threads, int* variant) {
  if (shape < 0 || shape >= kFusedShapes) return Fail(GDTP_ERROR_INVALID, "tile shape outside [0, 7)");
  const FusedShapeInfo& f = kFusedShape[shape];
  if (width) *width = f.ta;
  if (height) *height = f.tb;
  if (threads) *threads = f.threads;
  if (variant) *variant = f.variant;
  return GDTP_OK;
}

int gdtp_mixed_level_tile_shape(gdtp_system* s, int level, int shape, int* current) {
  if (!s || !s->ds || level < 0 || level >= s->ds->levels || !ValidShape(shape))
    return Fail(GDTP_ERROR_INVALID, "level or tile shape out of range");
  GDTP_TRY({
    MixedSystem& m = Mixed(s);  // the automatic shape depends on the level's size
    if (s->levelShape.empty()) s->levelShape.assign(s->ds->levels, s->tileShape);
    if (shape != -2 && shape != s->levelShape[level]) {
      s->levelShape[level] = shape;
      m.L[level].shape = shape == GDTP_TILE_SHAPE_AUTO ? AutoTileShape(m.L[level]) : shape;
      InvalidateGraph(m);
    }
    if (current) *current = m.L[level].shape;
    return GDTP_OK;
  })
}

int gdtp_mixed_configure(gdtp_system* s, int nuPre, int nuPost, int boundarySweeps, double omega, int useGraph) {
  if (!s || nuPre < 0 || nuPost < 0 || nuPre + nuPost == 0 || boundarySweeps < 0 || !(omega > 0 && omega <= 1.5))
    return Fail(GDTP_ERROR_INVALID, "invalid mixed-solver configuration");
  s->nuPre = nuPre, s->nuPost = nuPost, s->boundarySweeps = boundarySweeps, s->omega = omega, s->useGraph = useGraph ? 1 : 0;
  if (s->mixed) InvalidateGraph(*s->mixed);  // the captured step holds the previous schedule
  if (!s->ds) return GDTP_OK;
  GDTP_TRY({
    Mixed(s);  // build the mixed solver now (device set-up off the first solve's path)
    return GDTP_OK;
  })
}

int gdtp_mixed_acceleration(gdtp_system* s, int flexible_cg) {
  if (!s) return Fail(GDTP_ERROR_INVALID, "null system");
  s->krylov = flexible_cg ? 1 : 0;
  if (s->mixed) {
    s->mixed->krylov = s->krylov != 0;
    InvalidateGraph(*s->mixed);
  }
  return GDTP_OK;
}

int gdtp_mixed_cycles(gdtp_system* s, int cycles) {
  if (!s || cycles < 0 || cycles > 16) return Fail(GDTP_ERROR_INVALID, "V-cycles per iteration must be in [0, 16] (0: automatic)");
  s->cycles = cycles;
  if (s->mixed) {
    s->mixed->cycles = cycles;
    InvalidateGraph(*s->mixed);
  }
  return GDTP_OK;
}

int gdtp_mixed_deep_sweep(gdtp_system* s, int level, int ncols, void* stream) {
  GDTP_DEVICE_OP(level >= 0 && level + 1 < s->ds->levels, MixedDeepSweep(Mixed(s), level, ncols, st))
}

int gdtp_system_memory(const gdtp_system* s, size_t* refBytes, size_t* mixedBytes) {
  if (!s) return Fail(GDTP_ERROR_INVALID, "null system");
  if (refBytes) *refBytes = s->ds ? s->ds->bytes : 0;
  if (mixedBytes) *mixedBytes = s->mixed ? s->mixed->bytes : 0;
  return GDTP_OK;
}

int gdtp_solve_weighted(gdtp_system* s, double* x, const double* b, int ncols, const double* weights, double tol, int maxIters,
                        double* relRes, int* iterations, void* stream) {
  int rc = Ready(s, ncols);
  if (rc != GDTP_OK) return rc;
  if (!x || !b || !weights || !(tol > 0) || maxIters < 0) return Fail(GDTP_ERROR_INVALID, "invalid solve arguments");
  GDTP_TRY({
    std::vector<double> rel(ncols, 0.0);
    const int it = SolveMixed(*s->ds, Mixed(s), x, b, ncols, tol, maxIters, rel.data(), (hipStream_t)stream, s->useGraph != 0, weights);
    GDTP_HIP(hipGetLastError());
    bool ok = true;
    for (double r : rel) ok = ok && r <= tol;
    if (relRes) for (int k = 0; k < ncols; k++) relRes[k] = rel[k];
    if (iterations) *iterations = it;
    if (!ok) return Fail(GDTP_ERROR_NUMERIC, "solve did not reach the tolerance within max_iters");
    return GDTP_OK;
  })
}

int gdtp_geodesic_divergence(gdtp_system* s, const double* potential, double* rhs, int ncols, void* stream) {
  GDTP_DEVICE_OP(potential && rhs, {
    if (!s->vf) {
      auto d = std::make_unique<DevVectorField>();
      UploadVectorField(Samples(s->H), *d);
      s->vf = std::move(d);
    }
    VectorFieldDivergence(*s->vf, potential, ds.L[0].N, rhs, ds.L[0].N, ds.L[0].N, ncols, st);
  })
}

int gdtp_solve(gdtp_system* s, double* x, const double* b, int ncols, double tol, int maxIters, int mode, double* relRes, int* iterations,
               void* stream) {
  return gdtp_solve_yielding(s, x, b, ncols, tol, maxIters, mode, relRes, iterations, stream, nullptr, nullptr);
}

int gdtp_solve_yielding(gdtp_system* s, double* x, const double* b, int ncols, double tol, int maxIters, int mode, double* relRes,
                        int* iterations, void* stream, void (*between)(void*), void* user) {
  int rc = Ready(s, ncols);
  if (rc != GDTP_OK) return rc;
  const bool zeroGuess = (mode & GDTP_SOLVE_ZERO_GUESS) != 0;
  mode &= ~GDTP_SOLVE_ZERO_GUESS;
  if (!x || !b || !(tol > 0) || maxIters < 0 || (mode != GDTP_SOLVE_FP64_REFERENCE && mode != GDTP_SOLVE_MIXED))
    return Fail(GDTP_ERROR_INVALID, "invalid solve arguments");
  GDTP_TRY({
    hipStream_t st = (hipStream_t)stream;
    DeviceSystem& ds = *s->ds;
    std::vector<double> rel(ncols, 0.0);
    int it = 0;
    bool ok = false;
    if (mode == GDTP_SOLVE_MIXED) {
      MixedSystem& m = Mixed(s);
      it = SolveMixed(ds, m, x, b, ncols, tol, maxIters, rel.data(), st, s->useGraph != 0, nullptr, between, user, zeroGuess);
      ok = true;
      for (double r : rel) ok = ok && r <= tol;
    } else {
      if (!s->dNorms) {
        GDTP_HIP(hipMalloc(&s->dNorms, 2 * s->kcap * sizeof(double)));
        GDTP_HIP(hipHostMalloc((void**)&s->hNorms, 2 * s->kcap * sizeof(double), hipHostMallocDefault));
      }
      const int N = ds.L[0].N;
      if (zeroGuess) GDTP_HIP(hipMemsetAsync(x, 0, (size_t)N * ncols * sizeof(double), st));  // this loop is never captured
      SquaredNorms(b, N, ncols, N, s->dNorms + s->kcap, st);
      for (;; it++) {
        ResidualLevel(ds, 0, x, b, ds.L[0].r, ncols, N, st);
        SquaredNorms(ds.L[0].r, N, ncols, N, s->dNorms, st);
        GDTP_HIP(hipMemcpyAsync(s->hNorms, s->dNorms, 2 * s->kcap * sizeof(double), hipMemcpyDeviceToHost, st));
        GDTP_HIP(hipStreamSynchronize(st));
        ok = true;
        for (int k = 0; k < ncols; k++) {
          rel[k] = std::sqrt(s->hNorms[k]) / std::sqrt(std::max(s->hNorms[s->kcap + k], 1e-300));
          ok = ok && rel[k] <= tol;
        }
        if (ok || it >= maxIters) break;
        if (between && it > 0) between(user);
        VCycleReference(ds, x, b, ncols, N, st);
      }
    }
    GDTP_HIP(hipGetLastError());
    if (relRes) for (int k = 0; k < ncols; k++) relRes[k] = rel[k];
    if (iterations) *iterations = it;
    if (!ok) return Fail(GDTP_ERROR_NUMERIC, "solve did not reach the tolerance within max_iters");
    return GDTP_OK;
  })
}

// ---------------------------------------------------------------- stiffness modulation
int gdtp_modulation_create(const gdtp_hierarchy* h, const double* signal, int ncols, gdtp_modulation** out) {
  if (!h || !signal || !out || ncols < 1 || ncols > 1024) return Fail(GDTP_ERROR_INVALID, "invalid modulation arguments");
  GDTP_TRY({
    auto m = std::make_unique<gdtp_modulation>();
    FineOperator op;
    std::string err;
    if (!AssembleFineOperator(h->mesh, h->h, h->metric, op, err, signal, ncols, &m->host)) return Fail(GDTP_ERROR_GEOMETRY, err);
    const Level& L0 = h->h.levels[0];
    m->N = L0.numTexels;
    m->D = L0.numDeep;
    m->B = L0.numBoundary;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess) devices = 0;
    if (devices > 0) {
      UploadModulation(L0, m->host, m->d);
      m->onDevice = true;
      m->host = CellStiffnessRhs();
    }
    *out = m.release();
    return GDTP_OK;
  })
}

int gdtp_modulated_stiffness(const gdtp_modulation* m, const double* mask, double* out, void* stream) {
  if (!m || !mask || !out) return Fail(GDTP_ERROR_INVALID, "null argument");
  if (!m->onDevice) return Fail(GDTP_ERROR_DEVICE, "modulation was built without a GPU");
  GDTP_TRY({
    ModulatedStiffness(m->d, mask, out, (hipStream_t)stream);
    GDTP_HIP(hipGetLastError());
    return GDTP_OK;
  })
}

int gdtp_modulation_info(const gdtp_modulation* m, int* info) {
  if (!m || !info) return Fail(GDTP_ERROR_INVALID, "null argument");
  if (m->onDevice) info[0] = m->d.K, info[2] = m->d.numCells, info[3] = (int)m->d.nnz;
  else info[0] = m->host.K, info[2] = m->host.numCells, info[3] = (int)m->host.bCell.size();
  info[1] = m->N;
  return GDTP_OK;
}

template <class T>
static int DeviceArray(const T* p, size_t n, T* out, size_t* count) {
  if (!count) return Fail(GDTP_ERROR_INVALID, "null count");
  if (out) {
    if (*count < n) return Fail(GDTP_ERROR_INVALID, "output buffer too small");
    if (n) GDTP_TRY({ GDTP_HIP(hipMemcpy(out, p, n * sizeof(T), hipMemcpyDeviceToHost)); })
  }
  *count = n;
  return GDTP_OK;
}

int gdtp_modulation_array_f64(const gdtp_modulation* m, const char* name, double* out, size_t* count) {
  if (!m || !name) return Fail(GDTP_ERROR_INVALID, "null argument");
  const DevModulation& d = m->d;
  const std::string n(name);
  if (n == "deep") return m->onDevice ? DeviceArray(d.deep, (size_t)d.K * 4 * d.D, out, count) : Emit(m->host.deep, out, count);
  if (n == "boundaryVal") return m->onDevice ? DeviceArray(d.bVal, (size_t)d.K * d.nnz, out, count) : Emit(m->host.bVal, out, count);
  return Fail(GDTP_ERROR_INVALID, "unknown modulation array " + n);
}

int gdtp_modulation_array_i32(const gdtp_modulation* m, const char* name, int* out, size_t* count) {
  if (!m || !name) return Fail(GDTP_ERROR_INVALID, "null argument");
  const DevModulation& d = m->d;
  const std::string n(name);
  if (n == "deepCells") return m->onDevice ? DeviceArray(d.deepCells, 2 * (size_t)d.D, out, count) : Emit(m->host.deepCells, out, count);
  if (n == "boundaryRowptr") return m->onDevice ? DeviceArray(d.bRowptr, (size_t)d.B + 1, out, count) : Emit(m->host.bRowptr, out, count);
  if (n == "boundaryCell") return m->onDevice ? DeviceArray(d.bCell, (size_t)d.nnz, out, count) : Emit(m->host.bCell, out, count);
  return Fail(GDTP_ERROR_INVALID, "unknown modulation array " + n);
}

void gdtp_modulation_destroy(gdtp_modulation* m) { delete m; }
