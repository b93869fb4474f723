// Seeded synthetic textured meshes for the solver's workloads, a compact binary mesh file, and
// PLY input with per-corner texture coordinates (include/Src/SimpleMesh.h:118-160).
//
// The generated meshes: a unit cube (6 charts packed 3x2), a noisy icosphere (20 icosahedron-face
// charts) and a noisy torus (rectangular patches).  Every UV corner is kept at least 1e-3 texel away
// from texel-centre lines so that the reference's automatic grid-line jitter
// (include/Src/AtlasMesh.inl:36-55) never triggers.
#include "gdtp_mesh.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <sstream>

namespace gdtp {

static double NudgeOffCenterLine(double u, int res) {
  double t = u * double(res) - 0.5;
  double off = t - std::round(t);
  const double minOff = 1e-3;
  if (std::fabs(off) < minOff) t = std::round(t) + (off >= 0 ? minOff : -minOff);
  return (t + 0.5) / double(res);
}

static void NudgeUVs(TexturedMesh& m, int W, int H) {
  for (auto& p : m.uv) {
    p.x = NudgeOffCenterLine(p.x, W);
    p.y = NudgeOffCenterLine(p.y, H);
  }
}

// ---------------------------------------------------------------- cube
int MakeCube(int subdiv, int W, int H, TexturedMesh& m) {
  if (subdiv < 1 || W < 48 || H < 32) return 0;
  const int n = subdiv;
  std::map<std::array<int, 3>, int> lattice;
  auto vid = [&](int i, int j, int k) {
    std::array<int, 3> key{i, j, k};
    auto it = lattice.find(key);
    if (it != lattice.end()) return it->second;
    int id = (int)m.vertices.size();
    m.vertices.push_back(V3(double(i) / n - 0.5, double(j) / n - 0.5, double(k) / n - 0.5));
    lattice[key] = id;
    return id;
  };
  // (origin, U, V) in lattice units; U x V is the outward normal.
  const int faces[6][9] = {
      {0, 0, 0, 0, 1, 0, 1, 0, 0},  // z = 0, normal -z
      {0, 0, n, 1, 0, 0, 0, 1, 0},  // z = n, normal +z
      {0, 0, 0, 1, 0, 0, 0, 0, 1},  // y = 0, normal -y
      {0, n, 0, 0, 0, 1, 1, 0, 0},  // y = n, normal +y
      {0, 0, 0, 0, 0, 1, 0, 1, 0},  // x = 0, normal -x
      {n, 0, 0, 0, 1, 0, 0, 0, 1},  // x = n, normal +x
  };
  // Charts packed 3 x 2; each face is a square whose side is a whole number of texels per quad.
  // The u and v insets differ by 0.25 texel so that quad diagonals (x - y = const) pass a
  // quarter texel away from texel centres.
  const double cellW = double(W) / 3.0, cellH = double(H) / 2.0;
  const double s = n * std::floor((std::min(cellW, cellH) - 8.0) / n);
  if (s < n) return 0;  // fewer than one texel per quad
  for (int f = 0; f < 6; f++) {
    const int* F = faces[f];
    const int col = f % 3, row = f / 3;
    const double u0 = std::floor(col * cellW) + 3.3, v0 = std::floor(row * cellH) + 3.55;
    for (int b = 0; b < n; b++)
      for (int a = 0; a < n; a++) {
        int ids[2][2];
        V2 uvs[2][2];
        for (int db = 0; db < 2; db++)
          for (int da = 0; da < 2; da++) {
            int aa = a + da, bb = b + db;
            ids[db][da] = vid(F[0] + aa * F[3] + bb * F[6], F[1] + aa * F[4] + bb * F[7], F[2] + aa * F[5] + bb * F[8]);
            uvs[db][da] = V2((u0 + s * aa / n) / W, (v0 + s * bb / n) / H);
          }
        const int tri[2][3][2] = {{{0, 0}, {0, 1}, {1, 1}}, {{0, 0}, {1, 1}, {1, 0}}};  // (db,da)
        for (int t = 0; t < 2; t++)
          for (int k = 0; k < 3; k++) {
            m.triangles.push_back(ids[tri[t][k][0]][tri[t][k][1]]);
            m.uv.push_back(uvs[tri[t][k][0]][tri[t][k][1]]);
          }
      }
  }
  NudgeUVs(m, W, H);
  return 1;
}

// ---------------------------------------------------------------- smooth noise
// Sum of eight seeded plane waves on R^3, normalised to unit maximum amplitude.
struct SmoothNoise {
  V3 k[8];
  double phase[8], amp[8], norm = 1;
  explicit SmoothNoise(uint64_t seed) {
    SplitMix64 rng(seed);
    double s = 0;
    for (int i = 0; i < 8; i++) {
      double z = 2 * rng.uniform() - 1, t = 2 * M_PI * rng.uniform(), f = 1.5 + 2.5 * rng.uniform();
      double rr = std::sqrt(std::max(0.0, 1 - z * z));
      k[i] = V3(rr * std::cos(t), rr * std::sin(t), z) * f;
      phase[i] = 2 * M_PI * rng.uniform();
      amp[i] = 0.5 + rng.uniform();
      s += amp[i];
    }
    norm = 1.0 / s;
  }
  double operator()(V3 p) const {
    double v = 0;
    for (int i = 0; i < 8; i++) v += amp[i] * std::sin(dot(k[i], p) + phase[i]);
    return v * norm;
  }
};

// ---------------------------------------------------------------- icosphere
int MakeIcosphere(int levels, int W, int H, double amplitude, uint64_t seed, TexturedMesh& m) {
  if (levels < 0 || levels > 10) return 0;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double bv[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                            {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  int bf[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                   {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                   {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  V3 base[12];
  for (int i = 0; i < 12; i++) base[i] = V3(bv[i][0], bv[i][1], bv[i][2]);
  for (int f = 0; f < 20; f++) {  // make every base face outward-oriented
    V3 a = base[bf[f][0]], b = base[bf[f][1]], c = base[bf[f][2]];
    if (dot(cross(b - a, c - a), a + b + c) < 0) std::swap(bf[f][1], bf[f][2]);
  }
  const int n = 1 << levels;
  SmoothNoise noise(seed);
  // Shared vertices are keyed by their integer barycentric weights over the base vertices.
  std::map<std::array<int, 6>, int> vmap;
  auto vid = [&](int f, int i, int j) {
    int w[3] = {n - i - j, i, j};
    std::array<int, 6> key{-1, 0, -1, 0, -1, 0};
    int c = 0;
    int ids[3] = {bf[f][0], bf[f][1], bf[f][2]};
    // sort non-zero (id, weight) pairs by id
    std::vector<std::pair<int, int>> pw;
    for (int q = 0; q < 3; q++)
      if (w[q]) pw.push_back({ids[q], w[q]});
    std::sort(pw.begin(), pw.end());
    for (auto& p : pw) key[2 * c] = p.first, key[2 * c + 1] = p.second, c++;
    auto it = vmap.find(key);
    if (it != vmap.end()) return it->second;
    V3 p(0, 0, 0);
    for (auto& q : pw) p = p + base[q.first] * (double(q.second) / n);
    V3 u = p / norm(p);
    double r = 1.0 + amplitude * noise(u);
    int id = (int)m.vertices.size();
    m.vertices.push_back(u * r);
    vmap[key] = id;
    return id;
  };
  // Atlas layout: 4 rows of 5 equilateral triangles alternating up/down, side L texels.
  const double margin = 3.0, gap = 4.0;
  const double h_over_L = std::sqrt(3.0) / 2.0;
  double L = std::min((double(H) - 2 * margin - 3 * gap) / 4.0 / h_over_L, (double(W) - 2 * margin - 4 * gap) / 3.0);
  L = std::floor(L) - 0.37;
  const double hgt = L * h_over_L;
  for (int f = 0; f < 20; f++) {
    int row = f / 5, k = f % 5;
    double y0 = margin + row * (hgt + gap), sk = margin + k * (L / 2 + gap);
    V2 A, B, C;
    if (k % 2 == 0) {
      A = V2(sk, y0), B = V2(sk + L, y0), C = V2(sk + L / 2, y0 + hgt);
    } else {
      A = V2(sk + L, y0 + hgt), B = V2(sk, y0 + hgt), C = V2(sk + L / 2, y0);
    }
    auto uvAt = [&](int i, int j) {
      V2 p = A + (B - A) * (double(i) / n) + (C - A) * (double(j) / n);
      return V2(p.x / W, p.y / H);
    };
    for (int i = 0; i < n; i++)
      for (int j = 0; j + i < n; j++) {
        int a = vid(f, i, j), b = vid(f, i + 1, j), c = vid(f, i, j + 1);
        m.triangles.insert(m.triangles.end(), {a, b, c});
        m.uv.insert(m.uv.end(), {uvAt(i, j), uvAt(i + 1, j), uvAt(i, j + 1)});
        if (i + j < n - 1) {
          int d = vid(f, i + 1, j + 1);
          m.triangles.insert(m.triangles.end(), {b, d, c});
          m.uv.insert(m.uv.end(), {uvAt(i + 1, j), uvAt(i + 1, j + 1), uvAt(i, j + 1)});
        }
      }
  }
  NudgeUVs(m, W, H);
  return 1;
}

// ---------------------------------------------------------------- torus
int MakeTorus(int nu, int nv, int chartsU, int chartsV, double R, double r, double amplitude, uint64_t seed, int W,
              int H, TexturedMesh& m) {
  if (nu % chartsU || nv % chartsV) return 0;
  SmoothNoise noise(seed);
  m.vertices.resize((size_t)nu * nv);
  for (int j = 0; j < nv; j++)
    for (int i = 0; i < nu; i++) {
      double th = 2 * M_PI * i / nu, ph = 2 * M_PI * j / nv;
      V3 c(R * std::cos(th), R * std::sin(th), 0);
      V3 d(std::cos(ph) * std::cos(th), std::cos(ph) * std::sin(th), std::sin(ph));
      double rr = r * (1.0 + amplitude * noise(c + d * r));
      m.vertices[(size_t)j * nu + i] = c + d * rr;
    }
  // charts laid out on a grid of cols x rows cells; each chart is a (pu x pv) quad patch.
  const int pu = nu / chartsU, pv = nv / chartsV;
  int nCharts = chartsU * chartsV;
  int cols = 1;
  while (cols * cols < nCharts) cols++;
  // prefer a layout whose cell aspect matches the patch aspect (major direction wider)
  int rows = (nCharts + cols - 1) / cols;
  const double cellW = double(W) / cols, cellH = double(H) / rows;
  // whole texels per quad; u and v insets differ by 0.25 texel so diagonals miss texel centres
  const double sw = pu * std::floor((cellW - 8.0) / pu), sh = pv * std::floor((cellH - 8.0) / pv);
  if (sw <= 0 || sh <= 0) return 0;
  for (int cv = 0; cv < chartsV; cv++)
    for (int cu = 0; cu < chartsU; cu++) {
      int c = cv * chartsU + cu;
      double u0 = std::floor((c % cols) * cellW) + 3.3, v0 = std::floor((c / cols) * cellH) + 3.55;
      for (int b = 0; b < pv; b++)
        for (int a = 0; a < pu; a++) {
          int i = cu * pu + a, j = cv * pv + b;
          int id00 = j * nu + i, id10 = j * nu + (i + 1) % nu, id11 = ((j + 1) % nv) * nu + (i + 1) % nu,
              id01 = ((j + 1) % nv) * nu + i;
          V2 q00((u0 + sw * a / pu) / W, (v0 + sh * b / pv) / H), q10((u0 + sw * (a + 1) / pu) / W, (v0 + sh * b / pv) / H),
              q11((u0 + sw * (a + 1) / pu) / W, (v0 + sh * (b + 1) / pv) / H), q01((u0 + sw * a / pu) / W, (v0 + sh * (b + 1) / pv) / H);
          m.triangles.insert(m.triangles.end(), {id00, id10, id11, id00, id11, id01});
          m.uv.insert(m.uv.end(), {q00, q10, q11, q00, q11, q01});
        }
    }
  NudgeUVs(m, W, H);
  return 1;
}

// ---------------------------------------------------------------- texture
void MakeNoiseTexture(int W, int H, int channels, uint64_t seed, std::vector<float>& tex) {
  tex.resize((size_t)W * H * channels);
  SplitMix64 rng(seed);
  for (size_t i = 0; i < tex.size(); i++) tex[i] = float(rng.uniform());
}

// ---------------------------------------------------------------- file I/O
// Layout: "GDTPMSH1", int32 nV, int32 nT, f64 V[3nV], i32 T[3nT], f64 UV[6nT].
int WriteMesh(const char* path, const TexturedMesh& m) {
  FILE* f = fopen(path, "wb");
  if (!f) return 0;
  int nv = (int)m.vertices.size(), nt = m.numTriangles();
  fwrite("GDTPMSH1", 1, 8, f);
  fwrite(&nv, 4, 1, f);
  fwrite(&nt, 4, 1, f);
  fwrite(m.vertices.data(), sizeof(V3), nv, f);
  fwrite(m.triangles.data(), 4, 3 * (size_t)nt, f);
  fwrite(m.uv.data(), sizeof(V2), 3 * (size_t)nt, f);
  fclose(f);
  return 1;
}

int ReadMesh(const char* path, TexturedMesh& m) {
  FILE* f = fopen(path, "rb");
  if (!f) return 0;
  char magic[8];
  int nv, nt;
  bool ok = fread(magic, 1, 8, f) == 8 && !memcmp(magic, "GDTPMSH1", 8) && fread(&nv, 4, 1, f) == 1 && fread(&nt, 4, 1, f) == 1;
  if (ok) {
    m.vertices.resize(nv);
    m.triangles.resize(3 * (size_t)nt);
    m.uv.resize(3 * (size_t)nt);
    ok = fread(m.vertices.data(), sizeof(V3), nv, f) == (size_t)nv &&
         fread(m.triangles.data(), 4, 3 * (size_t)nt, f) == 3 * (size_t)nt &&
         fread(m.uv.data(), sizeof(V2), 3 * (size_t)nt, f) == 3 * (size_t)nt;
  }
  fclose(f);
  return ok ? 1 : 0;
}



// ---------------------------------------------------------------------------- PLY input
// Triangle meshes with per-corner texture coordinates in the layout the reference reads
// (SimpleMesh.h:118-160): element vertex with x, y, z; element face with a vertex_indices list
// and a texcoord list (u, v per corner).  ASCII and binary little-endian; other properties skipped.
namespace {

struct PlyProp {
  std::string name, type, countType;
  bool list = false;
};
struct PlyElement {
  std::string name;
  long count = 0;
  std::vector<PlyProp> props;
};

int TypeSize(const std::string& t) {
  if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
  if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
  if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
  if (t == "double" || t == "float64") return 8;
  return 0;
}

double ReadBinary(const unsigned char* p, const std::string& t) {
  if (t == "char" || t == "int8") return (double)*(const int8_t*)p;
  if (t == "uchar" || t == "uint8") return (double)*p;
  int16_t s16;
  uint16_t u16;
  int32_t s32;
  uint32_t u32;
  float f;
  double d;
  if (t == "short" || t == "int16") return memcpy(&s16, p, 2), (double)s16;
  if (t == "ushort" || t == "uint16") return memcpy(&u16, p, 2), (double)u16;
  if (t == "int" || t == "int32") return memcpy(&s32, p, 4), (double)s32;
  if (t == "uint" || t == "uint32") return memcpy(&u32, p, 4), (double)u32;
  if (t == "float" || t == "float32") return memcpy(&f, p, 4), (double)f;
  return memcpy(&d, p, 8), d;
}

}  // namespace

int ReadPly(const std::string& path, TexturedMesh& m, std::string& err) {
  FILE* fp = fopen(path.c_str(), "rb");
  if (!fp) return err = "cannot open " + path, 0;
  std::vector<unsigned char> data;
  {
    unsigned char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), fp)) > 0) data.insert(data.end(), buf, buf + n);
    fclose(fp);
  }
  // header
  size_t pos = 0;
  auto line = [&](std::string& out) {
    if (pos >= data.size()) return false;
    size_t e = pos;
    while (e < data.size() && data[e] != '\n') e++;
    out.assign((const char*)&data[pos], e - pos);
    if (!out.empty() && out.back() == '\r') out.pop_back();
    pos = e + 1;
    return true;
  };
  std::string ln;
  if (!line(ln) || ln != "ply") return err = path + ": not a PLY file", 0;
  bool binary = false;
  std::vector<PlyElement> elems;
  for (;;) {
    if (!line(ln)) return err = path + ": truncated header", 0;
    std::istringstream is(ln);
    std::string kw;
    is >> kw;
    if (kw == "format") {
      std::string f;
      is >> f;
      if (f == "binary_little_endian") binary = true;
      else if (f != "ascii") return err = path + ": unsupported PLY format " + f, 0;
    } else if (kw == "element") {
      PlyElement e;
      is >> e.name >> e.count;
      elems.push_back(e);
    } else if (kw == "property") {
      if (elems.empty()) return err = path + ": property before element", 0;
      PlyProp p;
      std::string t;
      is >> t;
      if (t == "list") p.list = true, is >> p.countType >> p.type >> p.name;
      else p.type = t, is >> p.name;
      if (!TypeSize(p.type) || (p.list && !TypeSize(p.countType))) return err = path + ": unknown property type", 0;
      elems.back().props.push_back(p);
    } else if (kw == "end_header") break;
  }
  std::vector<V3> verts;
  std::vector<int> tris;
  std::vector<V2> uv;
  std::istringstream ascii;
  if (!binary) ascii.str(std::string((const char*)data.data() + pos, data.size() - pos));
  auto scalar = [&](const std::string& t, double& v) -> bool {
    if (!binary) return (bool)(ascii >> v);
    const int sz = TypeSize(t);
    if (pos + sz > data.size()) return false;
    v = ReadBinary(&data[pos], t);
    pos += sz;
    return true;
  };
  for (const PlyElement& e : elems) {
    const bool isVertex = e.name == "vertex", isFace = e.name == "face";
    for (long r = 0; r < e.count; r++) {
      V3 p;
      std::vector<double> idx, tc;
      for (const PlyProp& pr : e.props) {
        if (pr.list) {
          double cnt;
          if (!scalar(pr.countType, cnt)) return err = path + ": truncated data", 0;
          std::vector<double> vals((size_t)cnt);
          for (auto& v : vals)
            if (!scalar(pr.type, v)) return err = path + ": truncated data", 0;
          if (isFace && (pr.name == "vertex_indices" || pr.name == "vertex_index")) idx = vals;
          if (isFace && pr.name == "texcoord") tc = vals;
        } else {
          double v;
          if (!scalar(pr.type, v)) return err = path + ": truncated data", 0;
          if (isVertex && pr.name == "x") p.x = v;
          if (isVertex && pr.name == "y") p.y = v;
          if (isVertex && pr.name == "z") p.z = v;
        }
      }
      if (isVertex) verts.push_back(p);
      if (isFace) {
        if (idx.size() != 3 || tc.size() != 6) return err = path + ": faces must be triangles with 6 texcoords", 0;
        for (int k = 0; k < 3; k++) tris.push_back((int)idx[k]), uv.push_back(V2(tc[2 * k], tc[2 * k + 1]));
      }
    }
  }
  for (int v : tris)
    if (v < 0 || v >= (int)verts.size()) return err = path + ": vertex index out of range", 0;
  if (tris.empty()) return err = path + ": no faces", 0;
  m.vertices = std::move(verts), m.triangles = std::move(tris), m.uv = std::move(uv);
  return 1;
}

}  // namespace gdtp
