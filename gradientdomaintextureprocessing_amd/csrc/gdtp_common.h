// Small geometry / utility types shared by the native host code of the texel-domain solver.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace gdtp {

struct V2 {
  double x = 0, y = 0;
  V2() {}
  V2(double a, double b) : x(a), y(b) {}
  double operator[](int i) const { return i ? y : x; }
  double& operator[](int i) { return i ? y : x; }
};
inline V2 operator+(V2 a, V2 b) { return V2(a.x + b.x, a.y + b.y); }
inline V2 operator-(V2 a, V2 b) { return V2(a.x - b.x, a.y - b.y); }
inline V2 operator*(V2 a, double s) { return V2(a.x * s, a.y * s); }
inline V2 operator*(double s, V2 a) { return V2(a.x * s, a.y * s); }
inline V2 operator/(V2 a, double s) { return V2(a.x / s, a.y / s); }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline double cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() {}
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, double s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, double s) { return V3(a.x / s, a.y / s, a.z / s); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// Row-major 2x2 matrix m[r][c]; m(r,c) in the usual sense.
struct M2 {
  double a[2][2] = {{0, 0}, {0, 0}};
  double& operator()(int r, int c) { return a[r][c]; }
  double operator()(int r, int c) const { return a[r][c]; }
  double det() const { return a[0][0] * a[1][1] - a[0][1] * a[1][0]; }
  M2 inv() const {
    M2 r;
    double d = det();
    r.a[0][0] = a[1][1] / d;
    r.a[1][1] = a[0][0] / d;
    r.a[0][1] = -a[0][1] / d;
    r.a[1][0] = -a[1][0] / d;
    return r;
  }
  M2 T() const {
    M2 r;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) r.a[i][j] = a[j][i];
    return r;
  }
};
inline M2 operator*(const M2& A, const M2& B) {
  M2 r;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) r.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j];
  return r;
}
inline V2 operator*(const M2& A, V2 v) { return V2(A.a[0][0] * v.x + A.a[0][1] * v.y, A.a[1][0] * v.x + A.a[1][1] * v.y); }

// SplitMix64: the seeded generator used for every synthetic input (meshes, textures, sources).
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // uniform in [0,1) with 53 random bits
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Triangle mesh with one UV per triangle corner (the layout of a PLY with per-face texcoord lists).
struct TexturedMesh {
  std::vector<V3> vertices;
  std::vector<int> triangles;  // 3 per triangle
  std::vector<V2> uv;          // 3 per triangle, corner order
  int numTriangles() const { return (int)triangles.size() / 3; }
};

}  // namespace gdtp
