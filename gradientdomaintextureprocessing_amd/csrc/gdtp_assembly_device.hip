// Interior-cell mass / stiffness assembly on the GPU (MassAndStiffness.h:5-405 for the cells whose
// four corners are interior texels; boundary cells and seams stay on the host).
//
// Every interior cell lies inside one chart, where the chart's triangles tile the parameter domain
// without overlap, so a cell's 4x4 Q1 matrices are the sum over the triangles covering it of the
// integrals over (triangle ∩ cell) — computed here per cell, triangles in ascending order, so the
// result does not depend on thread timing:
//   1. per triangle: count it into every interior cell of its bounding box     (KCellCount)
//   2. host: prefix sum of the counts
//   3. per triangle: append its index to those cells' lists                    (KCellFill)
//   4. per cell: sort the short list, clip each triangle to the cell, integrate with the degree-4
//      six-point rule exactly as the host does, write the cell's matrices      (KCellIntegrate)
//   5. per deep texel: gather the nine stencil entries from its four cells     (KDeepGather)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <stdexcept>
#include <vector>

#include "gdtp_assembly.h"
#include "gdtp_device.h"

namespace gdtp {

namespace {

struct DevChartGrid {
  double cellW, cellH;
  int width, height;  // texel grid; cells are (width-1) x (height-1)
  int mapOffset;      // into the concatenated per-chart interior-cell maps
  int cellBase;       // global interior-cell index of the chart's first interior cell
};

struct DevTriangle {
  double p[6];      // chart-local corner coordinates (x0, y0, x1, y1, x2, y2)
  double ginv[4];   // inverse metric, row-major
  double gscale;    // sqrt(det metric)
  int chart;
};

__constant__ double cQuadPos[6][2] = {{0.816847572980459, 0.0915762135097705}, {0.0915762135097705, 0.816847572980459},
                                      {0.0915762135097705, 0.0915762135097705}, {0.108103018168070, 0.445948490915965},
                                      {0.445948490915965, 0.108103018168070},  {0.445948490915965, 0.445948490915965}};
__constant__ double cQuadW[6] = {0.109951743655322, 0.109951743655322, 0.109951743655322,
                                 0.223381589678011, 0.223381589678011, 0.223381589678011};

// cell range of a triangle's bounding box, as the host computes it
__device__ __forceinline__ void CellRange(const DevTriangle& t, const DevChartGrid& g, int& i0, int& i1, int& j0, int& j1) {
  const double mnx = fmin(fmin(t.p[0], t.p[2]), t.p[4]), mny = fmin(fmin(t.p[1], t.p[3]), t.p[5]);
  const double mxx = fmax(fmax(t.p[0], t.p[2]), t.p[4]), mxy = fmax(fmax(t.p[1], t.p[3]), t.p[5]);
  i0 = (int)floor(fmax(mnx, 0.0) / g.cellW), j0 = (int)floor(fmax(mny, 0.0) / g.cellH);
  i1 = (int)ceil(fmin(mxx, 1.0) / g.cellW), j1 = (int)ceil(fmin(mxy, 1.0) / g.cellH);
  i0 = max(i0, 0), j0 = max(j0, 0), i1 = min(i1, g.width - 1), j1 = min(j1, g.height - 1);
}

__global__ void KCellCount(int nTri, const DevTriangle* __restrict__ tris, const DevChartGrid* __restrict__ charts,
                           const int* __restrict__ cellMap, int* counts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nTri) return;
  const DevTriangle tr = tris[t];
  const DevChartGrid g = charts[tr.chart];
  int i0, i1, j0, j1;
  CellRange(tr, g, i0, i1, j0, j1);
  for (int j = j0; j < j1; j++)
    for (int i = i0; i < i1; i++) {
      const int li = cellMap[g.mapOffset + j * (g.width - 1) + i];
      if (li >= 0) atomicAdd(&counts[g.cellBase + li], 1);
    }
}

__global__ void KCellFill(int nTri, const DevTriangle* __restrict__ tris, const DevChartGrid* __restrict__ charts,
                          const int* __restrict__ cellMap, const long* __restrict__ offsets, int* cursor, int* lists) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nTri) return;
  const DevTriangle tr = tris[t];
  const DevChartGrid g = charts[tr.chart];
  int i0, i1, j0, j1;
  CellRange(tr, g, i0, i1, j0, j1);
  for (int j = j0; j < j1; j++)
    for (int i = i0; i < i1; i++) {
      const int li = cellMap[g.mapOffset + j * (g.width - 1) + i];
      if (li < 0) continue;
      const int cell = g.cellBase + li;
      const int slot = atomicAdd(&cursor[cell], 1);
      lists[offsets[cell] + slot] = t;
    }
}

// Sutherland-Hodgman clip of a convex polygon against {p : dot(p, n) > offset} (the host's
// ClipConvex, same arithmetic); at most 7 vertices for a triangle clipped by four half-planes.
__device__ __forceinline__ int Clip(const double* px, const double* py, int n, double nx, double ny, double off, double* qx, double* qy) {
  if (n == 0) return 0;
  int m = 0;
  double ax = px[n - 1], ay = py[n - 1];
  double pl = ax * nx + ay * ny - off;
  bool pin = pl > 0;
  for (int k = 0; k < n; k++) {
    const double cx = px[k], cy = py[k];
    const double cl = cx * nx + cy * ny - off;
    const bool cin = cl > 0;
    if (cin != pin) {
      const double a = -cl / (pl - cl);
      qx[m] = ax * a + cx * (1.0 - a), qy[m] = ay * a + cy * (1.0 - a);
      m++;
    }
    if (cin) qx[m] = cx, qy[m] = cy, m++;
    ax = cx, ay = cy, pl = cl, pin = cin;
  }
  return m;
}

__device__ __forceinline__ double Q1(int k, double x, double y) {
  return k == 0 ? (1 - x) * (1 - y) : k == 1 ? x * (1 - y) : k == 2 ? x * y : (1 - x) * y;
}
__device__ __forceinline__ void Q1Grad(int k, double x, double y, double& gx, double& gy) {
  if (k == 0) gx = y - 1, gy = x - 1;
  else if (k == 1) gx = 1 - y, gy = -x;
  else if (k == 2) gx = y, gy = x;
  else gx = -y, gy = 1 - x;
}

__global__ __launch_bounds__(128) void KCellIntegrate(int nCells, const long* __restrict__ offsets, int* lists,
                                                      const DevTriangle* __restrict__ tris, const DevChartGrid* __restrict__ charts,
                                                      const int* __restrict__ cellChart, const int* __restrict__ cellIJ,
                                                      double* cellM, double* cellS) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nCells) return;
  const long b = offsets[c], e = offsets[c + 1];
  for (long x = b + 1; x < e; x++) {  // ascending triangle order (the lists are short)
    const int v = lists[x];
    long y = x - 1;
    while (y >= b && lists[y] > v) lists[y + 1] = lists[y], y--;
    lists[y + 1] = v;
  }
  const DevChartGrid g = charts[cellChart[c]];
  const int i = cellIJ[2 * c], j = cellIJ[2 * c + 1];
  const double cw = g.cellW, ch = g.cellH;
  double M[16], S[16];
  for (int k = 0; k < 16; k++) M[k] = 0.0, S[k] = 0.0;
  for (long x = b; x < e; x++) {
    const DevTriangle tr = tris[lists[x]];
    double ax[8], ay[8], bx[8], by[8];
    for (int k = 0; k < 3; k++) ax[k] = tr.p[2 * k], ay[k] = tr.p[2 * k + 1];
    int n = 3;
    n = Clip(ax, ay, n, 0, 1, j * ch, bx, by);
    n = Clip(bx, by, n, 0, -1, -(j + 1) * ch, ax, ay);
    n = Clip(ax, ay, n, 1, 0, i * cw, bx, by);
    n = Clip(bx, by, n, -1, 0, -(i + 1) * cw, ax, ay);
    for (int p = 2; p < n; p++) {  // fan from the first vertex
      const double d0x = ax[p - 1] - ax[0], d0y = ay[p - 1] - ay[0], d1x = ax[p] - ax[0], d1y = ay[p] - ay[0];
      const double area = fabs(d0x * d1y - d0y * d1x) * tr.gscale / 2.0;
      double val[4][6], gx[4][6], gy[4][6];
      for (int s = 0; s < 6; s++) {
        const double qx = ax[0] + d0x * cQuadPos[s][0] + d1x * cQuadPos[s][1];
        const double qy = ay[0] + d0y * cQuadPos[s][0] + d1y * cQuadPos[s][1];
        const double u = qx / cw - i, w = qy / ch - j;
        for (int k = 0; k < 4; k++) {
          val[k][s] = Q1(k, u, w);
          double hx, hy;
          Q1Grad(k, u, w, hx, hy);
          gx[k][s] = hx / cw, gy[k][s] = hy / ch;
        }
      }
      for (int k = 0; k < 4; k++)
        for (int l = 0; l < 4; l++) {
          double m = 0, st = 0;
          for (int s = 0; s < 6; s++) {
            m += val[k][s] * val[l][s] * cQuadW[s];
            const double tx = tr.ginv[0] * gx[k][s] + tr.ginv[1] * gy[k][s], ty = tr.ginv[2] * gx[k][s] + tr.ginv[3] * gy[k][s];
            st += (gx[l][s] * tx + gy[l][s] * ty) * cQuadW[s];
          }
          M[4 * l + k] += m * area;
          S[4 * l + k] += st * area;
        }
    }
  }
  for (int k = 0; k < 16; k++) cellM[16L * c + k] = M[k], cellS[16L * c + k] = S[k];
}

// Deep texel d: its four interior cells (i-1,j-1), (i,j-1), (i-1,j), (i,j) in that order, the
// texel at corner (2, 3, 1, 0) of each; stencil slot 3(dj+1)+(di+1) of neighbour (di,dj).
__global__ void KDeepGather(int D, const int* __restrict__ deepCells, const double* __restrict__ cellM, const double* __restrict__ cellS,
                            double* deepM, double* deepS) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D) return;
  const int offI[4] = {0, 1, 1, 0}, offJ[4] = {0, 0, 1, 1}, cornerOf[4] = {2, 3, 1, 0};
  double m[10], s[10];
  for (int k = 0; k < 10; k++) m[k] = 0.0, s[k] = 0.0;
  for (int slot = 0; slot < 4; slot++) {
    const long c = deepCells[4 * d + slot];
    const int k = cornerOf[slot];
    for (int l = 0; l < 4; l++) {
      const int n = (offJ[l] - offJ[k] + 1) * 3 + (offI[l] - offI[k] + 1);
      m[n] += cellM[16 * c + 4 * k + l];
      s[n] += cellS[16 * c + 4 * k + l];
    }
  }
  for (int k = 0; k < 10; k++) deepM[10L * d + k] = m[k], deepS[10L * d + k] = s[k];
}

__global__ void KGatherCells(int n, const int* __restrict__ ids, const double* __restrict__ cellM, const double* __restrict__ cellS,
                             double* outM, double* outS) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 16L * n) return;
  const long c = ids[t / 16];
  outM[t] = cellM[16 * c + t % 16], outS[t] = cellS[16 * c + t % 16];
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  explicit DevBuf(size_t n) {
    if (n) GDTP_HIP(hipMalloc(&p, n * sizeof(T)));
  }
  DevBuf(const std::vector<T>& v) : DevBuf(v.size()) {
    if (!v.empty()) GDTP_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

inline unsigned Blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void AssembleInteriorCellsOnDevice(const Hierarchy& h, const std::vector<std::vector<M2>>& metric, const std::vector<int>& wanted,
                                   std::vector<double>& deepMass, std::vector<double>& deepStiffness, std::vector<double>& wantedM,
                                   std::vector<double>& wantedS) {
  const Level& L0 = h.levels[0];
  const int nCharts = (int)L0.charts.size();
  std::vector<DevChartGrid> charts(nCharts);
  std::vector<int> cellMap, cellChart, cellIJ;
  int base = 0;
  for (int c = 0; c < nCharts; c++) {
    const GridChart& g = L0.charts[c];
    charts[c] = {g.cellW, g.cellH, g.width, g.height, (int)cellMap.size(), base};
    cellMap.insert(cellMap.end(), g.localInteriorCell.begin(), g.localInteriorCell.end());
    cellChart.resize(base + g.numInteriorCells, c);
    cellIJ.resize(2 * ((size_t)base + g.numInteriorCells));
    for (int j = 0; j + 1 < g.height; j++)
      for (int i = 0; i + 1 < g.width; i++) {
        const int li = g.localInteriorCell[(size_t)j * (g.width - 1) + i];
        if (li >= 0) cellIJ[2 * ((size_t)base + li)] = i, cellIJ[2 * ((size_t)base + li) + 1] = j;
      }
    base += g.numInteriorCells;
  }
  const int nCells = base;
  std::vector<DevTriangle> tris;
  for (int c = 0; c < nCharts; c++) {
    const GridChart& g = L0.charts[c];
    const AtlasChart& ac = h.atlas.charts[c];
    for (int t = 0; t < (int)ac.triangles.size() / 3; t++) {
      DevTriangle d;
      for (int k = 0; k < 3; k++) {
        const V2 p = ac.vertices[ac.triangles[3 * t + k]] - g.corner;
        d.p[2 * k] = p.x, d.p[2 * k + 1] = p.y;
      }
      const M2 Gi = metric[c][t].inv();
      d.ginv[0] = Gi(0, 0), d.ginv[1] = Gi(0, 1), d.ginv[2] = Gi(1, 0), d.ginv[3] = Gi(1, 1);
      d.gscale = std::sqrt(metric[c][t].det());
      d.chart = c;
      tris.push_back(d);
    }
  }
  // deep texels: their four cells (global interior-cell indices)
  std::vector<int> deepCells(4 * (size_t)L0.numDeep, -1);
  for (int c = 0; c < nCharts; c++) {
    const GridChart& g = L0.charts[c];
    auto cell = [&](int i, int j) { return charts[c].cellBase + g.localInteriorCell[(size_t)j * (g.width - 1) + i]; };
    for (int j = 1; j + 1 < g.height; j++)
      for (int i = 1; i + 1 < g.width; i++) {
        const int d = g.deep(i, j);
        if (d < 0) continue;
        const int cs[4] = {cell(i - 1, j - 1), cell(i, j - 1), cell(i - 1, j), cell(i, j)};
        for (int s = 0; s < 4; s++) {
          if (cs[s] < charts[c].cellBase) throw std::runtime_error("deep texel next to a non-interior cell");
          deepCells[4 * (size_t)d + s] = cs[s];
        }
      }
  }
  const int nTri = (int)tris.size();
  DevBuf<DevTriangle> dTris(tris);
  DevBuf<DevChartGrid> dCharts(charts);
  DevBuf<int> dMap(cellMap), dCellChart(cellChart), dCellIJ(cellIJ), dDeepCells(deepCells);
  DevBuf<int> dCounts(std::vector<int>(nCells, 0));
  if (nTri > 0) KCellCount<<<Blocks(nTri, 256), 256>>>(nTri, dTris.p, dCharts.p, dMap.p, dCounts.p);
  GDTP_HIP(hipGetLastError());
  std::vector<int> counts(nCells);
  if (nCells) GDTP_HIP(hipMemcpy(counts.data(), dCounts.p, nCells * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<long> offsets(nCells + 1, 0);
  for (int c = 0; c < nCells; c++) offsets[c + 1] = offsets[c] + counts[c];
  DevBuf<long> dOffsets(offsets);
  DevBuf<int> dCursor(std::vector<int>(nCells, 0));
  DevBuf<int> dLists((size_t)offsets[nCells]);
  if (nTri > 0) KCellFill<<<Blocks(nTri, 256), 256>>>(nTri, dTris.p, dCharts.p, dMap.p, dOffsets.p, dCursor.p, dLists.p);
  DevBuf<double> dCellM(16 * (size_t)nCells), dCellS(16 * (size_t)nCells);
  if (nCells > 0)
    KCellIntegrate<<<Blocks(nCells, 128), 128>>>(nCells, dOffsets.p, dLists.p, dTris.p, dCharts.p, dCellChart.p, dCellIJ.p, dCellM.p, dCellS.p);
  const int D = L0.numDeep;
  DevBuf<double> dDeepM(10 * (size_t)D), dDeepS(10 * (size_t)D);
  if (D > 0) KDeepGather<<<Blocks(D, 256), 256>>>(D, dDeepCells.p, dCellM.p, dCellS.p, dDeepM.p, dDeepS.p);
  const int nw = (int)wanted.size();
  DevBuf<int> dWanted(wanted);
  DevBuf<double> dWM(16 * (size_t)nw), dWS(16 * (size_t)nw);
  if (nw > 0) KGatherCells<<<Blocks(16L * nw, 256), 256>>>(nw, dWanted.p, dCellM.p, dCellS.p, dWM.p, dWS.p);
  GDTP_HIP(hipGetLastError());
  deepMass.resize(10 * (size_t)D), deepStiffness.resize(10 * (size_t)D);
  wantedM.resize(16 * (size_t)nw), wantedS.resize(16 * (size_t)nw);
  if (D) {
    GDTP_HIP(hipMemcpy(deepMass.data(), dDeepM.p, deepMass.size() * sizeof(double), hipMemcpyDeviceToHost));
    GDTP_HIP(hipMemcpy(deepStiffness.data(), dDeepS.p, deepStiffness.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (nw) {
    GDTP_HIP(hipMemcpy(wantedM.data(), dWM.p, wantedM.size() * sizeof(double), hipMemcpyDeviceToHost));
    GDTP_HIP(hipMemcpy(wantedS.data(), dWS.p, wantedS.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
}

}  // namespace gdtp
