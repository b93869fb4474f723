// Mixed-precision multi-RHS multigrid (the fast path).
//
// Outer loop in FP64 on the compact texel vectors (iterative refinement):
//     r = b - A x  (FP64 stencils, written as FP32 into the level-0 dense planes, norms in FP64)
//     e = V(r)     (one FP32 V-cycle of the correction equation A e = r, e = 0 initially)
//     x += e
// The FP32 V-cycle works on dense per-chart grids split into four colour planes by texel parity
// (q = (i&1) | (j&1)<<1; texel (i,j) lives at plane q, sub-grid (i>>1, j>>1)).  A four-colour
// Gauss-Seidel sweep then updates one plane per launch from fully coalesced neighbour planes.
// Boundary / seam texels are relaxed by restricted additive Schwarz: overlapping blocks of up to
// 64 boundary rows, each solved exactly through a precomputed inverse (seam coupling makes the
// boundary rows far from diagonally dominant, so pointwise smoothers stall); the coarsest level
// is solved exactly in FP64 (dense inverse when small, else LDL^T).
// Transfers are the reference's: 3x3 full weighting / the boundary restriction matrix, and
// bilinear prolongation (IterativeSolvers.inl:1009-1168), evaluated from chart geometry.
#include "gdtp_mixed.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <stdexcept>
#include <type_traits>
#include <unordered_map>

namespace gdtp {

static constexpr int kTA = 64, kTB = 4;  // tile of the sub-grid per workgroup: 64 x 4 (256 threads)
static constexpr int kDenseCoarseMax = 6000;  // coarsest levels up to this size use a dense inverse
static constexpr int kDotBlocks = 1024;       // partial sums per column of the flexible-CG dot products
static constexpr long kChartSweepMax = 2048;  // largest chart colour plane (texels) swept by one workgroup
static inline unsigned NBlk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ __forceinline__ long DenseIdx(const MixChart& c, long plane, int i, int j) {
  int q = (i & 1) | ((j & 1) << 1);
  return (long)q * plane + c.base2 + (long)(j >> 1) * c.w2 + (i >> 1);
}

// Dense vectors are addressed with 32-bit byte offsets (each column is under 2 GiB, checked at
// upload) from a uniform per-column base held in a buffer resource: one offset register per
// neighbour serves all columns of a chunk, which keeps the register count low while every load
// of the chunk is in flight.
__device__ __forceinline__ float Ld(const float* column, int bytes) {
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(column), 0, 0x7fffffff, 0x00020000);
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, bytes, 0, 0));
}
__device__ __forceinline__ int DenseByte(const MixChart& c, long plane, int i, int j) {
  return 4 * (((i & 1) | ((j & 1) << 1)) * (int)plane + c.base2 + (j >> 1) * c.w2 + (i >> 1));
}

// Tile (or block of rows) of workgroup b when the eight XCDs (workgroup b runs on XCD b % 8) each
// take a contiguous eighth of the n tiles, in order: XCD x owns tiles [x per + min(x, rem), ...),
// per + (x < rem) of them (a bijection on [0, n)).  Neighbouring tiles then share an L2, so the halo
// texels a tile reads from its neighbours are mostly L2 hits instead of HBM re-reads.
__device__ __forceinline__ int XcdTile(int b, int n) {
  const int x = b % 8, k = b / 8, per = n / 8, rem = n % 8;
  return x * per + min(x, rem) + k;
}

// Columns in chunks of up to four with the chunk width as a compile-time constant, so that a
// kernel can issue every load of a chunk before its first store (one memory round trip per chunk
// instead of one per column).
// kMax: the widest chunk (4, or 3 where a kernel holds many values per column and the register
// count of the widest chunk would halve its occupancy).
template <int kMax = 4, class F>
__device__ __forceinline__ void ForColumnChunks(int K, F&& f) {
  for (int c0 = 0; c0 < K; c0 += kMax) {
    switch (min(kMax, K - c0)) {
      case 4: if constexpr (kMax >= 4) f(c0, std::integral_constant<int, 4>()); break;
      case 3: f(c0, std::integral_constant<int, 3>()); break;
      case 2: f(c0, std::integral_constant<int, 2>()); break;
      default: f(c0, std::integral_constant<int, 1>()); break;
    }
  }
}

// ------------------------------------------------------------------ deep four-colour Gauss-Seidel
// Update of deep texel (i, j) of plane q (sub-grid (a, b)): e = c4 rhs - sum_{k!=4} c_k e_k.  The
// neighbours lie in the other three planes, so every load of a chunk of columns is issued before
// its stores (few memory round trips per texel: small levels are latency bound).
__device__ __forceinline__ float Half(unsigned bits) { return (float)__builtin_bit_cast(_Float16, (unsigned short)bits); }

// The eight FP16 off-diagonal ratios of a packed stencil (slot 4 unused).
template <bool kHalf>
__device__ __forceinline__ void Unpack(const uint4& w, float (&cf)[9]) {
  cf[0] = Half(w.x & 0xffff), cf[1] = Half(w.x >> 16), cf[2] = Half(w.y & 0xffff), cf[3] = Half(w.y >> 16), cf[4] = 0.0f;
  cf[5] = Half(w.z & 0xffff), cf[6] = Half(w.z >> 16), cf[7] = Half(w.w & 0xffff), cf[8] = Half(w.w >> 16);
}

// The Gauss-Seidel value of one texel and column from its eight neighbours v (slot 4 unused):
// c4 rhs - sum_{k!=4} c_k v_k, with the FP16 stencil's row-sum correction - corr * sum_k v_k.  The
// fused-multiply-adds are explicit so that every kernel that calls this evaluates it identically.
template <bool kHalf>
__device__ __forceinline__ float GsValue(const float (&cf)[9], const float (&v)[9], float c4, float corr, float r) {
  float sum = 0.0f, m = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; k++)
    if (k != 4) {
      sum = __builtin_fmaf(cf[k], v[k], sum);
      if constexpr (kHalf) m += v[k];
    }
  const float e = __builtin_fmaf(c4, r, -sum);
  if constexpr (kHalf) return __builtin_fmaf(-corr, m, e);
  return e;
}

// The smoother's stencil of dense texel idx: c4 (0 off deep texels), the FP16 stencil's row-sum
// correction and the eight off-diagonal ratios.
template <bool kHalf>
__device__ __forceinline__ float GsStencil(long idx, long Nd, const float* __restrict__ st, const float2* __restrict__ stc,
                                           const uint4* __restrict__ sth, float& corr, float (&cf)[9]) {
  corr = 0.0f;
  if constexpr (kHalf) {
    const float2 cd = stc[idx];
    if (cd.x == 0.0f) return 0.0f;
    corr = 0.125f * cd.y;
    const uint4 w = sth[idx];
    cf[0] = Half(w.x & 0xffff), cf[1] = Half(w.x >> 16), cf[2] = Half(w.y & 0xffff), cf[3] = Half(w.y >> 16), cf[4] = 0.0f;
    cf[5] = Half(w.z & 0xffff), cf[6] = Half(w.z >> 16), cf[7] = Half(w.w & 0xffff), cf[8] = Half(w.w >> 16);
    return cd.x;
  } else {
    const float c4 = st[4 * Nd + idx];
    if (c4 == 0.0f) return 0.0f;
#pragma unroll
    for (int k = 0; k < 9; k++) cf[k] = k == 4 ? 0.0f : st[k * Nd + idx];
    return c4;
  }
}

// kHalf: the stencil as FP16 ratios with an FP32 row-sum correction (stc, sth), else the FP32
// stencil st.
template <bool kHalf>
__device__ __forceinline__ void DeepGsTexel(const MixChart& c, long plane, long Nd, int q, int a, int b, const float* __restrict__ st,
                                            const float2* __restrict__ stc, const uint4* __restrict__ sth, float* e,
                                            const float* __restrict__ rhs, int K) {
  const int i = 2 * a + (q & 1), j = 2 * b + (q >> 1);
  if (i >= c.w || j >= c.h) return;
  const long idx = (long)q * plane + c.base2 + (long)b * c.w2 + a;
  float corr, cf[9];
  const float c4 = GsStencil<kHalf>(idx, Nd, st, stc, sth, corr, cf);
  if (c4 == 0.0f) return;
  int nb[9];
#pragma unroll
  for (int k = 0; k < 9; k++) nb[k] = DenseByte(c, plane, i + k % 3 - 1, j + k / 3 - 1);
  ForColumnChunks(K, [&](int c0, auto kc) {
    constexpr int KC = decltype(kc)::value;
    float v[KC][9], r[KC];
#pragma unroll
    for (int ch = 0; ch < KC; ch++) {
#pragma unroll
      for (int k = 0; k < 9; k++)
        if (k != 4) v[ch][k] = Ld(e + (c0 + ch) * Nd, nb[k]);
      r[ch] = Ld(rhs + (c0 + ch) * Nd, nb[4]);
    }
#pragma unroll
    for (int ch = 0; ch < KC; ch++) e[(long)(c0 + ch) * Nd + idx] = GsValue<kHalf>(cf, v[ch], c4, corr, r[ch]);
  });
}

// One colour plane of a tiled level: wave w of the workgroup takes row w of its 64 x 4 tile, whose
// deep texels it reads from a host-built bit mask (one uniform 64-bit load per wave) instead of
// from the stencil, so that a deep texel's stencil, neighbour and rhs loads all go out together
// (one memory round trip instead of two).  Tiles in XCD-contiguous order (XcdTile).
template <bool kHalf>
__global__ __launch_bounds__(256) void KMixDeepGs(int q, int nTiles, const int* __restrict__ tiles,
                                                  const unsigned long long* __restrict__ deepMask, const MixChart* __restrict__ charts,
                                                  long plane, long Nd, const float* __restrict__ st, const float2* __restrict__ stc,
                                                  const uint4* __restrict__ sth, float* e, const float* __restrict__ rhs, int K) {
  const int tile = XcdTile(blockIdx.x, nTiles), row = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTA), tx = threadIdx.x % kTA;
  const unsigned long long m = deepMask[((long)tile * 4 + q) * kTB + row];
  if (!((m >> tx) & 1ull)) return;
  const int* T = tiles + 3 * tile;
  const MixChart c = charts[T[0]];
  const int a = T[1] + tx, b = T[2] + row, i = 2 * a + (q & 1), j = 2 * b + (q >> 1);
  const int idx = (int)((long)q * plane + c.base2 + (long)b * c.w2 + a);
  float c4, corr = 0.0f, cf[9];
  if constexpr (kHalf) {
    const float2 cd = stc[idx];
    const uint4 w = sth[idx];
    c4 = cd.x, corr = 0.125f * cd.y;
    Unpack<true>(w, cf);
  } else {
    c4 = st[4 * Nd + idx];
#pragma unroll
    for (int k = 0; k < 9; k++) cf[k] = k == 4 ? 0.0f : st[k * Nd + idx];
  }
  int nb[9];
#pragma unroll
  for (int k = 0; k < 9; k++) nb[k] = DenseByte(c, plane, i + k % 3 - 1, j + k / 3 - 1);
  ForColumnChunks(K, [&](int c0, auto kc) {
    constexpr int KC = decltype(kc)::value;
    float v[KC][9], r[KC];
#pragma unroll
    for (int ch = 0; ch < KC; ch++) {
#pragma unroll
      for (int k = 0; k < 9; k++)
        if (k != 4) v[ch][k] = Ld(e + (c0 + ch) * Nd, nb[k]);
      r[ch] = Ld(rhs + (c0 + ch) * Nd, nb[4]);
    }
#pragma unroll
    for (int ch = 0; ch < KC; ch++) e[(long)(c0 + ch) * Nd + idx] = GsValue<kHalf>(cf, v[ch], c4, corr, r[ch]);
  });
}

// DeepGsTexel for the small levels swept one chart per workgroup, which latency bounds: a texel
// whose 3x3 neighbourhood lies inside the chart (a test on its position alone; every deep texel
// passes it) issues its stencil, neighbour and rhs loads together and only then learns from c4
// whether it is deep, one memory round trip instead of two.  (The other texels cannot be deep.)
template <bool kHalf>
__device__ __forceinline__ void ChartGsTexel(const MixChart& c, long plane, long Nd, int q, int a, int b, const float* __restrict__ st,
                                             const float2* __restrict__ stc, const uint4* __restrict__ sth, float* e,
                                             const float* __restrict__ rhs, int K) {
  const int i = 2 * a + (q & 1), j = 2 * b + (q >> 1);
  if (i < 1 || j < 1 || i + 1 >= c.w || j + 1 >= c.h) return;
  const int idx = (int)((long)q * plane + c.base2 + (long)b * c.w2 + a);
  float c4, corr = 0.0f, cf[9];
  if constexpr (kHalf) {
    const float2 cd = stc[idx];
    const uint4 w = sth[idx];
    c4 = cd.x, corr = 0.125f * cd.y;
    Unpack<true>(w, cf);
  } else {
    c4 = st[4 * Nd + idx];
#pragma unroll
    for (int k = 0; k < 9; k++) cf[k] = k == 4 ? 0.0f : st[k * Nd + idx];
  }
  int nb[9];
#pragma unroll
  for (int k = 0; k < 9; k++) nb[k] = DenseByte(c, plane, i + k % 3 - 1, j + k / 3 - 1);
  ForColumnChunks(K, [&](int c0, auto kc) {
    constexpr int KC = decltype(kc)::value;
    float v[KC][9], r[KC];
#pragma unroll
    for (int ch = 0; ch < KC; ch++) {
#pragma unroll
      for (int k = 0; k < 9; k++)
        if (k != 4) v[ch][k] = Ld(e + (c0 + ch) * Nd, nb[k]);
      r[ch] = Ld(rhs + (c0 + ch) * Nd, nb[4]);
    }
    if (c4 == 0.0f) return;  // not deep: nothing to update
#pragma unroll
    for (int ch = 0; ch < KC; ch++) e[(long)(c0 + ch) * Nd + idx] = GsValue<kHalf>(cf, v[ch], c4, corr, r[ch]);
  });
}

// The same four-colour sweep with one 1024-thread workgroup per chart, colours in order (post:
// 3..0) separated by barriers: deep stencils never reach across charts, so this is exactly the
// four launches of KMixDeepGs in one, for levels whose charts are small enough that launch
// latency, not memory, bounds a sweep.  `sweeps` consecutive sweeps (post-smoothing runs its sweeps
// back to back) in one launch.  All threads reach every barrier.
template <bool kHalf>
__global__ __launch_bounds__(1024) void KMixChartGs(int post, int sweeps, const MixChart* __restrict__ charts, long plane, long Nd,
                                                    const float* __restrict__ st, const float2* __restrict__ stc,
                                                    const uint4* __restrict__ sth, float* e, const float* __restrict__ rhs, int K) {
  const MixChart c = charts[blockIdx.x];
  const int aw = (c.w + 1) / 2, n = aw * c.h2;
  for (int sw = 0; sw < sweeps; sw++)
    for (int qq = 0; qq < 4; qq++) {
      const int q = post ? 3 - qq : qq;
      for (int t = threadIdx.x; t < n; t += blockDim.x) ChartGsTexel<kHalf>(c, plane, Nd, q, t % aw, t / aw, st, stc, sth, e, rhs, K);
      __syncthreads();
    }
}

// ------------------------------------------------------------------ colour pairs
// The colours of a sweep in pairs, (0, 1) then (2, 3) (post: (3, 2) then (1, 0)).  The second
// colour qb = qa ^ 1 of a pair reads the first only at its two horizontal neighbours, which lie in
// the same sub-grid row, and neither colour reads a value the other pair writes in this launch.  A
// workgroup owns one sub-grid row of one chart, so a pair takes one barrier and no halo: the first
// colour's row (new values at deep texels, current ones at the other active texels) goes to LDS,
// the second colour reads its two horizontal neighbours there and its other six from global
// memory.  Every plane is then read twice per sweep instead of three times, and the iterate is
// bitwise that of one launch per colour (same neighbour values, GsValue).  A row is the span
// [a0, a1) of sub-grid columns holding an active texel of any plane (the horizontal neighbours of a
// deep texel are active, so they lie in it).  Where the level's widest span has at most
// kPairThreadsMax texels, spans are packed whole into workgroups of NT = 256, 512 or 1024 threads, 64
// texels per wave (P = 1); wider ones take a workgroup of kPairThreadsMax threads each, thread t
// their texels a0 + t, a0 + t + NT, ... (P of them).  Workgroups in XCD-contiguous order
// (XcdTile), so the rows above and below share an L2.  (Measured at C2 level 0: one row per
// workgroup, 221 us per sweep with 256 threads of two texels each and the second colour's loads
// issued after the barrier, 210 us with 512 threads of one texel each, against 187 us for one
// launch per colour: the rows of triangular charts leave half the waves without texels.)
// One texel of a colour pair: its stencil, then its neighbour values (the second colour's two
// horizontal neighbours come from LDS later) and rhs; a texel that is not deep takes its current
// value (zero outside the chart), which the first colour puts into LDS.
template <bool kHalf, int KC>
struct PairTexel {
  float c4, corr, cf[9], v[KC][9], r[KC];
  int idx;
  bool act, deep;
};
// deep: 1 / 0 where the caller knows whether the texel is deep (its stencil and the other loads then
// go out together), -1 where the stencil tells (c4 != 0)
template <bool kHalf, int KC>
__device__ __forceinline__ void PairStencil(PairTexel<kHalf, KC>& T, const MixChart& c, long plane, long Nd, int q, int a, int a1, int b,
                                            int deep, const float* __restrict__ st, const float2* __restrict__ stc,
                                            const uint4* __restrict__ sth) {
  const int i = 2 * a + (q & 1), j = 2 * b + (q >> 1);
  T.act = a < a1 && i < c.w && j < c.h;
  T.idx = T.act ? (int)((long)q * plane + c.base2 + (long)b * c.w2 + a) : 0;
  T.corr = 0.0f;
  if (deep < 0) {
    T.c4 = T.act ? GsStencil<kHalf>(T.idx, Nd, st, stc, sth, T.corr, T.cf) : 0.0f;
    T.deep = T.c4 != 0.0f;
  } else {
    T.deep = deep != 0;
    T.c4 = 0.0f;
    if (T.deep) {
      if constexpr (kHalf) {
        const float2 cd = stc[T.idx];
        const uint4 w = sth[T.idx];
        T.c4 = cd.x, T.corr = 0.125f * cd.y;
        Unpack<true>(w, T.cf);
      } else {
        T.c4 = st[4 * Nd + T.idx];
#pragma unroll
        for (int k = 0; k < 9; k++) T.cf[k] = k == 4 ? 0.0f : st[k * Nd + T.idx];
      }
    }
  }
}
template <bool kHalf, int KC>
__device__ __forceinline__ void PairValues(PairTexel<kHalf, KC>& T, bool second, const MixChart& c, long plane, long Nd, int q, int a,
                                           int b, const float* e, const float* __restrict__ rhs) {
  if (T.deep) {
    const int i = 2 * a + (q & 1), j = 2 * b + (q >> 1);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int dx = k % 3 - 1, dy = k / 3 - 1;
      if (k == 4 || (second && dy == 0)) continue;
      const int nb = DenseByte(c, plane, i + dx, j + dy);
#pragma unroll
      for (int ch = 0; ch < KC; ch++) T.v[ch][k] = Ld(e + (long)ch * Nd, nb);
    }
#pragma unroll
    for (int ch = 0; ch < KC; ch++) T.r[ch] = Ld(rhs + (long)ch * Nd, 4 * T.idx);
  } else if (!second) {
#pragma unroll
    for (int ch = 0; ch < KC; ch++) T.v[ch][4] = T.act ? Ld(e + (long)ch * Nd, 4 * T.idx) : 0.0f;
  }
}
// The texel's new value (deep texels) written to e; the first colour's row value into LDS slot t;
// the second colour's horizontal neighbours read from LDS slots lo, lo + 1 first.
template <bool kHalf, int KC, int S>
__device__ __forceinline__ void PairUpdate(PairTexel<kHalf, KC>& T, bool second, int q, int t, float (&sm)[KC][S], float* e, long Nd) {
  if (second && T.deep) {
    const int lo = (q & 1) ? t : t - 1;  // span columns of i - 1 and i + 1 in the first colour's plane
#pragma unroll
    for (int ch = 0; ch < KC; ch++) T.v[ch][3] = sm[ch][lo], T.v[ch][5] = sm[ch][lo + 1];
  }
#pragma unroll
  for (int ch = 0; ch < KC; ch++) {
    if (T.deep) {
      const float val = GsValue<kHalf>(T.cf, T.v[ch], T.c4, T.corr, T.r[ch]);
      e[(long)ch * Nd + T.idx] = val;
      if (!second) sm[ch][t] = val;
    } else if (!second) {
      sm[ch][t] = T.v[ch][4];
    }
  }
}

template <bool kHalf, int KC, int NT, int P>
__global__ __launch_bounds__(NT) void KMixPairGs(int qa, const int* __restrict__ rows, const unsigned long long* __restrict__ rowMask,
                                                           int nRows, const MixChart* __restrict__ charts, long plane, long Nd,
                                                           const float* __restrict__ st, const float2* __restrict__ stc,
                                                           const uint4* __restrict__ sth, float* e, const float* __restrict__ rhs) {
  __shared__ float sm[KC][P * NT];
  const int qb = qa ^ 1;
  if constexpr (P == 1) {
    // packed rows: wave w takes the 64 texels [R.z, R.z + 64) of row (chart R.x, R.y), clipped to
    // the row's span end R.w; a row's waves are consecutive, so its texels have consecutive LDS
    // slots.  One texel per thread and colour; both colours' stencils are loaded together.
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64), slot = XcdTile(blockIdx.x, nRows) * (NT / 64) + w;
    const int4 R = reinterpret_cast<const int4*>(rows)[slot];
    const MixChart c = charts[R.x];
    const int b = R.y, a1 = R.w, t = threadIdx.x, a = R.z + t % 64;
    // the wave's deep texels in each plane (bit = lane): known before any stencil is loaded
    const int da = (int)((rowMask[4 * slot + qa] >> (t % 64)) & 1ull), db = (int)((rowMask[4 * slot + qb] >> (t % 64)) & 1ull);
    PairTexel<kHalf, KC> A, B;
    PairStencil(A, c, plane, Nd, qa, a, a1, b, da, st, stc, sth);
    PairStencil(B, c, plane, Nd, qb, a, a1, b, db, st, stc, sth);
    PairValues(A, false, c, plane, Nd, qa, a, b, e, rhs);
    // the second colour's neighbour loads go out once the first colour is done (its registers
    // free) and overlap the barrier wait: 60 registers, eight waves per SIMD (issuing them with
    // the first colour's took 78, six: C2 level 1 40.4 against 38.0 us per sweep)
    PairUpdate(A, false, qa, t, sm, e, Nd);
    PairValues(B, true, c, plane, Nd, qb, a, b, e, rhs);
    __syncthreads();
    PairUpdate(B, true, qb, t, sm, e, Nd);
  } else {  // one row per workgroup: span [R.z, R.w) of row (chart R.x, R.y)
    const int4 R = reinterpret_cast<const int4*>(rows)[XcdTile(blockIdx.x, nRows)];
    const MixChart c = charts[R.x];
    const int b = R.y, a0 = R.z, a1 = R.w;
    for (int p = 0; p < P; p++) {
      const int t = (int)threadIdx.x + NT * p, a = a0 + t;
      PairTexel<kHalf, KC> A;
      PairStencil(A, c, plane, Nd, qa, a, a1, b, -1, st, stc, sth);
      PairValues(A, false, c, plane, Nd, qa, a, b, e, rhs);
      PairUpdate(A, false, qa, t, sm, e, Nd);
    }
    __syncthreads();
    for (int p = 0; p < P; p++) {
      const int t = (int)threadIdx.x + NT * p, a = a0 + t;
      PairTexel<kHalf, KC> B;
      PairStencil(B, c, plane, Nd, qb, a, a1, b, -1, st, stc, sth);
      PairValues(B, true, c, plane, Nd, qb, a, b, e, rhs);
      PairUpdate(B, true, qb, t, sm, e, Nd);
    }
  }
}

__device__ __forceinline__ int FloorHalf(int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }

// The bilinear prolongation of the coarse column E (dense coarse planes) at fine texel
// (2a + (q & 1), 2b + (q >> 1)): the mean of its four (repeated) coarse parents, with the parents
// and their order of summation as in KMixProlongAdd, so adding it gives bitwise the same value.
__device__ __forceinline__ float ProlongTerm(const MixChart& f, const MixChart& c, long cplane, const float* E, int a, int b, int q) {
  const int ui0 = 2 * a - f.cofx, uj0 = 2 * b - f.cofy;
  const int ci0 = FloorHalf(ui0) + c.cofx, cj0 = FloorHalf(uj0) + c.cofy;
  const bool ei = (ui0 & 1) == 0, ej = (uj0 & 1) == 0;
  const int qi = q & 1, qj = q >> 1;
  const int bx = ei ? 0 : qi, by = ej ? 0 : qj;
  const bool ox = ei ? qi == 1 : qi == 0, oy = ej ? qj == 1 : qj == 0;
  const int need = ((1 | (ox ? 2 : 0)) * (1 | (oy ? 4 : 0))) << (bx + 2 * by);
  float V[4];
#pragma unroll
  for (int n = 0; n < 4; n++) V[n] = (need >> n) & 1 ? Ld(E, DenseByte(c, cplane, ci0 + (n & 1), cj0 + (n >> 1))) : 0.0f;
  const bool sx = !ei && qi == 1, sy = !ej && qj == 1;
  const float b0 = sx ? V[1] : V[0], b1 = sx ? V[3] : V[2];
  const float v00 = sy ? b1 : b0;
  const float v10 = ox ? (sy ? V[3] : V[1]) : v00;
  const float v01 = oy ? b1 : v00;
  const float v11 = (ox && oy) ? V[3] : (ox ? v10 : v01);
  return (v00 + v10 + v01 + v11) * 0.25f;
}

// Prolongation operands of the fused sweep that starts post-smoothing (null ce: no prolongation).
struct FusedProlong {
  const MixChart* cch = nullptr;
  long cplane = 0, cNd = 0;
  const float* ce = nullptr;
};

// ------------------------------------------------------------------ fused four-colour sweep
// All four colours of one sweep in one launch.  A workgroup owns a TA x TB tile of the sub-grid
// (all four planes) and stages the window [-2, TA+2) x [-1, TB+1) around it in LDS: the old
// values of the planes coloured after the first, then colour by colour the new values.  A colour
// is computed on its tile and, redundantly, on the ring of neighbouring-tile texels that a later
// colour of this tile reads (FusedRegion), so no workgroup waits for another.  The sweep reads e
// from eIn and writes every active texel of its tile to eOut (deep texels: new values; boundary
// texels: copied): neighbouring workgroups still read the old values from eIn, and the result is
// bitwise that of the four single-colour launches (same arithmetic, GsValue).  The caller swaps
// the two buffers afterwards.
struct FusedBox {
  int lx, hx, ly, hy;  // region [lx, TA + hx) x [ly, TB + hy) relative to the tile
};
// Region of plane q for the colour order 0,1,2,3 (pre-smoothing) or 3,2,1,0 (post): the plane's
// texels whose new value this tile's later colours read, closed over every later colour.
__host__ __device__ constexpr FusedBox FusedRegion(bool post, int q) {
  return post ? (q == 3 ? FusedBox{-2, 1, -1, 0} : q == 2 ? FusedBox{-1, 1, -1, 0} : q == 1 ? FusedBox{-1, 0, 0, 0} : FusedBox{0, 0, 0, 0})
              : (q == 0 ? FusedBox{-1, 2, 0, 1} : q == 1 ? FusedBox{-1, 1, 0, 1} : q == 2 ? FusedBox{0, 1, 0, 0} : FusedBox{0, 0, 0, 0});
}

// The operands one thread fetches for its texels of one colour's region (at most kIt of them):
// stencil (FP16 ratios still packed, or FP32), rhs and, for the first colour, the texel's old
// value where it is not deep.
struct Stencil9 {
  float v[9];
};
template <bool kHalf, bool kFirst, int KC, int kIt>
struct FusedOperands {
  float c4[kIt], corr[kIt], r[kIt][KC], old[kFirst ? kIt : 1][KC];
  typename std::conditional<kHalf, uint4, Stencil9>::type sten[kIt];
  int idx[kIt], w[kIt], state[kIt];  // state 0: outside the chart or region, 1: deep, 2: active not deep, 3: inactive
};

// Thread t's texels of the region [LX, TA + HX) x [LY, TB + HY) of plane Q: t, t + NT, ...
template <bool kHalf, bool kFirst, bool kProlong, int TA, int TB, int NT, int KC, int Q, int LX, int HX, int LY, int HY, int kIt>
__device__ __forceinline__ void FusedFetch(const MixChart& c, int a0, int b0, long plane, long Nd, const float* __restrict__ st,
                                           const float2* __restrict__ stc, const uint4* __restrict__ sth,
                                           const int8_t* __restrict__ type, const float* __restrict__ eIn, const float* __restrict__ rhs,
                                           const FusedProlong& pr, const MixChart& cc, FusedOperands<kHalf, kFirst, KC, kIt>& op) {
  constexpr int WX = TA + 4;
  constexpr int RW = TA + HX - LX, RN = RW * (TB + HY - LY);
#pragma unroll
  for (int it = 0; it < kIt; it++) {
    const int t = (int)threadIdx.x + NT * it;
    const int x = LX + t % RW, y = LY + t / RW;
    const int a = a0 + x, b = b0 + y, i = 2 * a + (Q & 1), j = 2 * b + (Q >> 1);
    const bool in = t < RN && a >= 0 && b >= 0 && i < c.w && j < c.h;
    const int idx = in ? (int)((long)Q * plane + c.base2 + (long)b * c.w2 + a) : 0;
    op.w[it] = (y + 1) * WX + (x + 2);
    op.idx[it] = idx;
    float c4 = 0.0f;
    if (in) {
      if constexpr (kHalf) {
        const float2 cd = stc[idx];
        c4 = cd.x, op.corr[it] = 0.125f * cd.y;
        if (c4 != 0.0f) op.sten[it] = sth[idx];
      } else {
        c4 = st[4 * Nd + idx], op.corr[it] = 0.0f;
        if (c4 != 0.0f) {
#pragma unroll
          for (int k = 0; k < 9; k++) op.sten[it].v[k] = k == 4 ? 0.0f : st[k * Nd + idx];
        }
      }
    }
    op.c4[it] = c4;
    const bool deep = in && c4 != 0.0f;
#pragma unroll
    for (int ch = 0; ch < KC; ch++) op.r[it][ch] = deep ? rhs[(long)ch * Nd + idx] : 0.0f;
    const int tp = in && !deep ? (int)type[idx] : -1;
    op.state[it] = !in ? 0 : (deep ? 1 : (tp >= 0 ? 2 : 3));
    if constexpr (kFirst) {
#pragma unroll
      for (int ch = 0; ch < KC; ch++) {
        float v = in && !deep ? eIn[(long)ch * Nd + idx] : 0.0f;
        if constexpr (kProlong)
          if (in && !deep && tp >= 0) v = v + ProlongTerm(c, cc, pr.cplane, pr.ce + (long)ch * pr.cNd, a, b, Q);
        op.old[it][ch] = v;
      }
    }
  }
}

// Colour Q on its region from the fetched operands: deep texels take the Gauss-Seidel value from
// their neighbours in LDS; the first colour's other texels put their old value into LDS; the
// tile's own active texels are written to eOut.
template <bool kHalf, bool kFirst, int TA, int TB, int NT, int KC, int Q, int LX, int HX, int LY, int HY, int kIt>
__device__ __forceinline__ void FusedApply(float* __restrict__ eOut, long Nd, float* sm, const FusedOperands<kHalf, kFirst, KC, kIt>& op) {
  constexpr int WX = TA + 4, WN = WX * (TB + 2);
  constexpr int RW = TA + HX - LX;
#pragma unroll
  for (int it = 0; it < kIt; it++) {
    const int st = op.state[it];
    const int t = (int)threadIdx.x + NT * it;
    const int x = LX + t % RW, y = LY + t / RW;
    const bool own = x >= 0 && x < TA && y >= 0 && y < TB;
    const int w = op.w[it];
    if (st == 0) {  // outside the chart (or past the region): the first colour zeroes its LDS slot
      if constexpr (kFirst) {
        if (t < RW * (TB + HY - LY)) {
#pragma unroll
          for (int ch = 0; ch < KC; ch++) sm[(ch * 4 + Q) * WN + w] = 0.0f;
        }
      }
      continue;
    }
    if (st == 1) {
      float cf[9];
      if constexpr (kHalf) Unpack<true>(op.sten[it], cf);
      else {
#pragma unroll
        for (int k = 0; k < 9; k++) cf[k] = op.sten[it].v[k];
      }
#pragma unroll
      for (int ch = 0; ch < KC; ch++) {
        float v[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
          if (k == 4) {
            v[k] = 0.0f;
            continue;
          }
          // neighbour (i + di, j + dj): plane p, sub-grid offset (da, db) (floor division by 2)
          const int ni = (Q & 1) + k % 3 - 1, nj = (Q >> 1) + k / 3 - 1;
          const int p = (ni & 1) | ((nj & 1) << 1), da = ni >> 1, db = nj >> 1;
          v[k] = sm[(ch * 4 + p) * WN + w + db * WX + da];
        }
        const float e = GsValue<kHalf>(cf, v, op.c4[it], op.corr[it], op.r[it][ch]);
        sm[(ch * 4 + Q) * WN + w] = e;
        if (own) eOut[(long)ch * Nd + op.idx[it]] = e;
      }
      continue;
    }
    if constexpr (kFirst) {
#pragma unroll
      for (int ch = 0; ch < KC; ch++) sm[(ch * 4 + Q) * WN + w] = op.old[it][ch];
    }
    if (own && st == 2) {
#pragma unroll
      for (int ch = 0; ch < KC; ch++) eOut[(long)ch * Nd + op.idx[it]] = sm[(ch * 4 + Q) * WN + w];
    }
  }
}

template <bool kPost, int TA, int TB, int NT, int S>
struct FusedPhase {
  static constexpr int Q = kPost ? 3 - S : S;
  static constexpr FusedBox R = FusedRegion(kPost, Q);
  static constexpr int RN = (TA + R.hx - R.lx) * (TB + R.hy - R.ly);
  static constexpr int kIt = (RN + NT - 1) / NT;
};

#define GDTP_FUSED_FETCH(S, OP)                                                                                                       \
  FusedFetch<kHalf, S == 0, kProlong, TA, TB, NT, KC, FusedPhase<kPost, TA, TB, NT, S>::Q, FusedPhase<kPost, TA, TB, NT, S>::R.lx,                \
             FusedPhase<kPost, TA, TB, NT, S>::R.hx, FusedPhase<kPost, TA, TB, NT, S>::R.ly, FusedPhase<kPost, TA, TB, NT, S>::R.hy,              \
             FusedPhase<kPost, TA, TB, NT, S>::kIt>(c, a0, b0, plane, Nd, st, stc, sth, type, eIn, rhs, pr, cc, OP)
#define GDTP_FUSED_APPLY(S, OP)                                                                                                       \
  FusedApply<kHalf, S == 0, TA, TB, NT, KC, FusedPhase<kPost, TA, TB, NT, S>::Q, FusedPhase<kPost, TA, TB, NT, S>::R.lx,                          \
             FusedPhase<kPost, TA, TB, NT, S>::R.hx, FusedPhase<kPost, TA, TB, NT, S>::R.ly, FusedPhase<kPost, TA, TB, NT, S>::R.hy,              \
             FusedPhase<kPost, TA, TB, NT, S>::kIt>(eOut, Nd, sm, OP)

// kProlong (with kPost): the sweep starts post-smoothing and first adds the prolongation of the
// coarse correction to every active texel it reads (what KMixProlongAdd would have written).
// kAhead: each colour's global loads (stencil, rhs, the first colour's old values) are issued right
// before the colour is computed (kFusedPerColour), one colour ahead (kFusedOneAhead: the next
// colour's operands are in flight while the current one is computed from LDS, so each barrier waits
// on at most one colour's compute), or all at the start (kFusedAllAhead: one memory round trip per
// tile).  NT threads per workgroup; kXcd: XCD-contiguous tile order (XcdTile).
template <bool kHalf, bool kPost, bool kProlong, int TA, int TB, int NT, int KC, int kAhead, bool kXcd>
__global__ __launch_bounds__(NT) void KMixDeepGsFused(const int* __restrict__ tiles, const MixChart* __restrict__ charts, long plane,
                                                      long Nd, const float* __restrict__ st, const float2* __restrict__ stc,
                                                      const uint4* __restrict__ sth, const int8_t* __restrict__ type,
                                                      const float* __restrict__ eIn, float* __restrict__ eOut,
                                                      const float* __restrict__ rhs, FusedProlong pr) {
  constexpr int WX = TA + 4, WN = WX * (TB + 2), kLoads = (3 * WN + NT - 1) / NT;
  __shared__ float sm[KC * 4 * WN];
  const int* T = tiles + 3 * (kXcd ? XcdTile(blockIdx.x, gridDim.x) : (int)blockIdx.x);
  const MixChart c = charts[T[0]];
  MixChart cc{};
  if constexpr (kProlong) cc = pr.cch[T[0]];
  const int a0 = T[1], b0 = T[2];
  FusedOperands<kHalf, true, KC, FusedPhase<kPost, TA, TB, NT, 0>::kIt> op0;
  FusedOperands<kHalf, false, KC, FusedPhase<kPost, TA, TB, NT, 1>::kIt> op1;
  FusedOperands<kHalf, false, KC, FusedPhase<kPost, TA, TB, NT, 2>::kIt> op2;
  FusedOperands<kHalf, false, KC, FusedPhase<kPost, TA, TB, NT, 3>::kIt> op3;
  if constexpr (kAhead != kFusedPerColour) GDTP_FUSED_FETCH(0, op0);
  if constexpr (kAhead == kFusedAllAhead) {
    GDTP_FUSED_FETCH(1, op1);
    GDTP_FUSED_FETCH(2, op2);
    GDTP_FUSED_FETCH(3, op3);
  }
  // old values of the three planes coloured after the first (pre: 1, 2, 3; post: 2, 1, 0)
  float v[kLoads][KC];
#pragma unroll
  for (int it = 0; it < kLoads; it++) {
    const int t = (int)threadIdx.x + NT * it;
    const int q = kPost ? 2 - t / WN : 1 + t / WN, w = t % WN;
    const int a = a0 + w % WX - 2, b = b0 + w / WX - 1, i = 2 * a + (q & 1), j = 2 * b + (q >> 1);
#pragma unroll
    for (int ch = 0; ch < KC; ch++) v[it][ch] = 0.0f;
    if (t < 3 * WN && a >= 0 && b >= 0 && i < c.w && j < c.h) {
      const long idx = (long)q * plane + c.base2 + (long)b * c.w2 + a;
#pragma unroll
      for (int ch = 0; ch < KC; ch++) v[it][ch] = eIn[(long)ch * Nd + idx];
      if constexpr (kProlong) {
        if (type[idx] >= 0) {
#pragma unroll
          for (int ch = 0; ch < KC; ch++) v[it][ch] = v[it][ch] + ProlongTerm(c, cc, pr.cplane, pr.ce + (long)ch * pr.cNd, a, b, q);
        }
      }
    }
  }
  // one colour ahead: the second colour's loads go out before the window is stored (and waited for)
  if constexpr (kAhead == kFusedOneAhead) GDTP_FUSED_FETCH(1, op1);
#pragma unroll
  for (int it = 0; it < kLoads; it++) {
    const int t = (int)threadIdx.x + NT * it;
    if (t >= 3 * WN) continue;
    const int q = kPost ? 2 - t / WN : 1 + t / WN, w = t % WN;
#pragma unroll
    for (int ch = 0; ch < KC; ch++) sm[(ch * 4 + q) * WN + w] = v[it][ch];
  }
  if constexpr (kAhead == kFusedAllAhead) {
    __syncthreads();
    GDTP_FUSED_APPLY(0, op0);
    __syncthreads();
    GDTP_FUSED_APPLY(1, op1);
    __syncthreads();
    GDTP_FUSED_APPLY(2, op2);
    __syncthreads();
    GDTP_FUSED_APPLY(3, op3);
  } else if constexpr (kAhead == kFusedOneAhead) {
    __syncthreads();
    GDTP_FUSED_APPLY(0, op0);
    GDTP_FUSED_FETCH(2, op2);
    __syncthreads();
    GDTP_FUSED_APPLY(1, op1);
    GDTP_FUSED_FETCH(3, op3);
    __syncthreads();
    GDTP_FUSED_APPLY(2, op2);
    __syncthreads();
    GDTP_FUSED_APPLY(3, op3);
  } else {
    __syncthreads();
    GDTP_FUSED_FETCH(0, op0);
    GDTP_FUSED_APPLY(0, op0);
    __syncthreads();
    GDTP_FUSED_FETCH(1, op1);
    GDTP_FUSED_APPLY(1, op1);
    __syncthreads();
    GDTP_FUSED_FETCH(2, op2);
    GDTP_FUSED_APPLY(2, op2);
    __syncthreads();
    GDTP_FUSED_FETCH(3, op3);
    GDTP_FUSED_APPLY(3, op3);
  }
}
#undef GDTP_FUSED_FETCH
#undef GDTP_FUSED_APPLY

// res = rhs - A e at deep texels (raw A: the symmetric five-word stencil, a4 the diagonal).  A
// workgroup covers its tile in all four colour planes, so the planes' shared neighbourhoods of e are
// fetched once rather than once per plane; its deep texels come from the tile's bit masks (one
// uniform load per wave and plane), so a texel's stencil, neighbour and rhs loads go out together.
// (Measured and not kept: a thread owning the 2x2 block of texels of one sub-grid position, its 4x4
// window of e and all four stencils loaded before any is used: one memory round trip per thread,
// but 138 registers and 232 instead of 221 us at C2; the tile's window of e in all four planes
// staged in LDS by row-contiguous loads, seven instead of 36 loads per column per thread: 220 us,
// as the HBM bytes, not the load count, bound the launch.)
// kChunk: columns whose loads are in flight together (2 on large levels: 50 instead of 79
// registers, eight waves per SIMD instead of six, level 0 at C2 219 against 221 us; 4 on small
// levels, which are latency-bound and pay for a second round trip).
template <int kChunk>
__device__ __forceinline__ void ResidualDeepTile(int vb, const int* __restrict__ tiles, const unsigned long long* __restrict__ deepMask,
                                                 const MixChart* __restrict__ charts, long plane, long Nd, const float* __restrict__ sym,
                                                 const float* __restrict__ e, const float* __restrict__ rhs, float* res, int K, int zeroQ) {
  const int row = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTA), tx = threadIdx.x % kTA;
  const unsigned long long* mk = deepMask + (long)vb * 4 * kTB + row;
  const int* T = tiles + 3 * vb;
  const MixChart c = charts[T[0]];
  const int a = T[1] + tx, b = T[2] + row;
  for (int q = 0; q < 4; q++) {
    if (!((mk[q * kTB] >> tx) & 1ull)) continue;
    const int i = 2 * a + (q & 1), j = 2 * b + (q >> 1);
    const long idx = (long)q * plane + c.base2 + (long)b * c.w2 + a;
    if (q == zeroQ) {  // relaxed last by the FP32 smoother: its deep rows hold exactly, so the residual is zero
      for (int ch = 0; ch < K; ch++) res[(long)ch * Nd + idx] = 0.0f;
      continue;
    }
    const float a4 = sym[4 * Nd + idx];
    int nb[9];
#pragma unroll
    for (int k = 0; k < 9; k++) nb[k] = DenseByte(c, plane, i + k % 3 - 1, j + k / 3 - 1);
    // own lower couplings and a4; upper coupling k from slot 8-k of that neighbour
    const float cf[9] = {sym[idx], sym[Nd + idx], sym[2 * Nd + idx], sym[3 * Nd + idx], a4,
                         Ld(sym + 3 * Nd, nb[5]), Ld(sym + 2 * Nd, nb[6]), Ld(sym + Nd, nb[7]), Ld(sym, nb[8])};
    ForColumnChunks<kChunk>(K, [&](int c0, auto kc) {
      constexpr int KC = decltype(kc)::value;
      float v[KC][9], r[KC];
#pragma unroll
      for (int ch = 0; ch < KC; ch++) {
#pragma unroll
        for (int k = 0; k < 9; k++) v[ch][k] = Ld(e + (c0 + ch) * Nd, nb[k]);
        r[ch] = Ld(rhs + (c0 + ch) * Nd, nb[4]);
      }
#pragma unroll
      for (int ch = 0; ch < KC; ch++) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 9; k++) s += cf[k] * v[ch][k];
        res[(long)(c0 + ch) * Nd + idx] = r[ch] - s;
      }
    });
  }
}

// ------------------------------------------------------------------ boundary rows
// Boundary rows are sparse (about 16 entries); a group of kRowLanes lanes works on one row so that
// the row's entries are read coalesced, once for up to four columns, and summed by shuffles.
static constexpr int kRowLanes = 16;    // FP32 boundary rows (measured best of 8 / 16)
static constexpr int kRowLanes64 = 8;   // FP64 level-0 boundary rows (M and S entries; 52 vs 73 us at 16)
template <class T, int Lanes = kRowLanes>
__device__ __forceinline__ T RowSum(T v) {
  for (int o = Lanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, Lanes);
  return v;
}
template <class T>
__device__ __forceinline__ T Pick(const T (&a)[4], int q) {
  return q == 0 ? a[0] : q == 1 ? a[1] : q == 2 ? a[2] : a[3];
}

// Boundary-row residual rhs - A e (FP32).  out: compact [K][B] when `compact`, else the dense
// vector [K][Nd] at the row's texel.  A row entry is one 64-bit word (dense column, FP32 value
// bits): one load per entry instead of two.
__device__ __forceinline__ void BoundaryResidualRows(int vb, int B, const int* __restrict__ bdIdx, const int* __restrict__ rp,
                                                     const int2* __restrict__ ent, const float* __restrict__ dg, const float* __restrict__ e,
                                                     const float* __restrict__ rhs, float* out, int compact, long Nd, int K) {
  const int r = (int)(((long)vb * blockDim.x + threadIdx.x) / kRowLanes), lane = threadIdx.x % kRowLanes;
  if (r >= B) return;  // whole row groups leave together
  const int k0 = rp[r], k1 = rp[r + 1];
  const long d = bdIdx[r];
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = min(4, K - c0);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = k0 + lane; k < k1; k += kRowLanes) {
      const int2 w = ent[k];
      const float v = __int_as_float(w.y);
      const long c = w.x;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q < kc) acc[q] += v * e[(long)(c0 + q) * Nd + c];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = RowSum(acc[q]);
    if (lane < kc) {
      const long o = (long)(c0 + lane) * Nd + d;
      const float rv = rhs[o] - (dg[r] * e[o] + Pick(acc, lane));
      if (compact) out[(long)(c0 + lane) * B + r] = rv;
      else out[o] = rv;
    }
  }
}

__global__ __launch_bounds__(256) void KMixBoundaryResidual(int B, const int* __restrict__ bdIdx, const int* __restrict__ rp,
                                                            const int2* __restrict__ ent, const float* __restrict__ dg,
                                                            const float* __restrict__ e, const float* __restrict__ rhs, float* out,
                                                            int compact, long Nd, int K) {
  BoundaryResidualRows(blockIdx.x, B, bdIdx, rp, ent, dg, e, rhs, out, compact, Nd, K);
}

static constexpr int kResidualWideTiles = 1024;  // levels with more 64 x 4 tiles load two columns at a time
// The residual of a level after pre-smoothing in one launch: blocks [0, nTiles) take the deep tiles,
// the others the boundary rows (disjoint outputs, so the two parts overlap on the device).
template <int kChunk>
__global__ __launch_bounds__(256) void KMixResidual(int nTiles, const int* __restrict__ tiles, const unsigned long long* __restrict__ deepMask,
                                                    const MixChart* __restrict__ charts, long plane,
                                                    long Nd, const float* __restrict__ sym, const float* __restrict__ e,
                                                    const float* __restrict__ rhs, float* res, int K, int zeroQ, int B,
                                                    const int* __restrict__ bdIdx, const int* __restrict__ rp, const int2* __restrict__ ent,
                                                    const float* __restrict__ dg) {
  if ((int)blockIdx.x < nTiles)
    ResidualDeepTile<kChunk>(XcdTile(blockIdx.x, nTiles), tiles, deepMask, charts, plane, Nd, sym, e, rhs, res, K, zeroQ);
  else BoundaryResidualRows(blockIdx.x - nTiles, B, bdIdx, rp, ent, dg, e, rhs, res, 0, Nd, K);
}

// Restricted additive Schwarz over the boundary rows: one workgroup of W waves per block, the
// waves splitting the extended rows; lane i of every wave owns core row i (< 64) and accumulates
// sum_j Minv[i][j] * res[ext_j] for up to four columns at a time, the block's residual staged in
// LDS.  Cores of different blocks are disjoint.  W = 4 on levels with at least
// kSchwarzManyBlocks blocks, where four times as many blocks are resident (C2 level 0: 40 against
// 45 us with 16 waves); W = 16 elsewhere, where a block's own latency bounds the step (level 3,
// 366 blocks: 10 us with 16 waves, 19 us with 4).
// TM: the inverse's entries as FP32 (float, one extended row per load) or as FP16 quads (uint2:
// extended rows 4s .. 4s + 3 of a core row in one 64-bit word, so a lane's load brings four
// entries), scaled per core row to |entry| <= 1 with the row's FP32 scale in rowScale.  The load
// count, not the bytes, bounds the step: FP16 with one entry per load saved nothing, two per load
// took C2 level 0 from 39.2 to 31.9 us, four per load C2 from 12.75 to 12.65 ms per solve.  The
// step is a smoother, so the FP16 rounding (2^-11 of the row's largest entry) perturbs its
// correction, not the residuals that drive the refinement.  Levels that smooth with FP16 stencil
// ratios (screened operators) and more than one block per CU hold FP16 quads beside the FP32
// inverses; plain refinement reads the quads, flexible CG the FP32 inverses (with FP16 entries,
// batched and single geodesic distances at 2048^2 drifted apart by 0.7 % of the largest distance
// at rel-res 1e-6, against under 0.1 % with FP32).  kPre counts loads (entries or quads).
static constexpr int kBlockCore = 64, kBlockExtMax = 384, kSchwarzWaves = 16;
static constexpr int kSchwarzPrefetchBlocks = 256;  // levels with at most this many blocks (one per CU) prefetch every entry
static constexpr int kSchwarzWavesLarge = 4;        // waves per block on levels with many blocks (more blocks resident)
static constexpr int kSchwarzManyBlocks = 1024;     // from this many blocks on
__device__ __forceinline__ void Entries(const float* M, long k, float (&m)[1]) { m[0] = M[k]; }
__device__ __forceinline__ void Entries(const uint2* M, long k, float (&m)[4]) {
  const uint2 w = M[k];
  m[0] = Half(w.x & 0xffff), m[1] = Half(w.x >> 16), m[2] = Half(w.y & 0xffff), m[3] = Half(w.y >> 16);
}
template <int W, int kPre, class TM>
__global__ __launch_bounds__(64 * W) void KMixSchwarz(const int* __restrict__ corePtr, const int* __restrict__ core,
                                                                  const int* __restrict__ extPtr, const int* __restrict__ ext,
                                                                  const long* __restrict__ matPtr, const TM* __restrict__ mat,
                                                                  const float* __restrict__ rowScale,
                                                                  const float* __restrict__ resB, const int* __restrict__ bdIdx,
                                                                  const float* __restrict__ rhs, int B, float* e, long Nd, int K,
                                                                  float omega) {
  constexpr bool kScaled = !std::is_same<TM, float>::value;
  constexpr int kJ = kScaled ? (int)sizeof(TM) / 2 : 1;  // extended rows per load
  static_assert(kBlockExtMax % kJ == 0, "padded extended rows fit the staged residual");
  __shared__ float R[4 * kBlockExtMax];
  __shared__ float P[W][4][kBlockCore];
  const int b = blockIdx.x, c0 = corePtr[b], nc = corePtr[b + 1] - c0, x0 = extPtr[b], nx = extPtr[b + 1] - x0;
  const TM* M = mat + matPtr[b];
  const int i = threadIdx.x % 64, part = threadIdx.x / 64;
  // The first kPre of this lane's loads of the block inverse (row i, load slots part, part + W,
  // ..., each kJ extended rows) are issued before the residual is staged, so they are in flight with it;
  // the rest stream in the loop.  All of them (kPre = kPerLane) on levels with at most one block
  // per CU, where latency bounds the step; none elsewhere (C2 level 0: 44 us; 61 us prefetching
  // all, 58 us prefetching 8 at the same occupancy).
  constexpr int kPerLane = kBlockExtMax / (kJ * W);
  static_assert(kPre >= 0 && kPre <= kPerLane, "prefetch depth");
  float mv[kPre > 0 ? kPre : 1][kJ];
#pragma unroll
  for (int ss = 0; ss < kPre; ss++) {
    const int sl = part + W * ss;
    if (i < nc && kJ * sl < nx) Entries(M, (long)sl * nc + i, mv[ss]);
    else
#pragma unroll
      for (int t = 0; t < kJ; t++) mv[ss][t] = 0.0f;
  }
  const int nxPad = (nx + kJ - 1) / kJ * kJ;  // a quad's rows past nx read a zero residual
  for (int ch0 = 0; ch0 < K; ch0 += 4) {
    const int kc = min(4, K - ch0);
    for (int j = threadIdx.x; j < nxPad; j += 64 * W)
      for (int q = 0; q < kc; q++)  // resB null: e is zero on this level, the residual is the rhs
        R[q * kBlockExtMax + j] = j >= nx ? 0.0f : resB ? resB[(long)(ch0 + q) * B + ext[x0 + j]] : rhs[(long)(ch0 + q) * Nd + bdIdx[ext[x0 + j]]];
    __syncthreads();
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (i < nc) {
#pragma unroll
      for (int ss = 0; ss < kPre; ss++) {
        const int j = kJ * (part + W * ss);
        if (j >= nx) break;
#pragma unroll
        for (int t = 0; t < kJ; t++)
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (q < kc) acc[q] += mv[ss][t] * R[q * kBlockExtMax + j + t];
      }
      for (int sl = part + W * kPre; kJ * sl < nx; sl += W) {
        float m[kJ];
        Entries(M, (long)sl * nc + i, m);
        const int j = kJ * sl;
#pragma unroll
        for (int t = 0; t < kJ; t++)
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (q < kc) acc[q] += m[t] * R[q * kBlockExtMax + j + t];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) P[part][q][i] = acc[q];
    __syncthreads();
    if (threadIdx.x < 4 * 64) {  // wave q sums column q over the parts
      const int q = threadIdx.x / 64;
      if (q < kc && i < nc) {
        float v = 0.0f;
        for (int p = 0; p < W; p++) v += P[p][q][i];
        if constexpr (kScaled) v *= rowScale[c0 + i];
        e[(long)(ch0 + q) * Nd + core[c0 + i]] += omega * v;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ transfers
// Coarse deep rhs: 3x3 full weighting of the fine residual around the fine texel 2c (all four
// coarse colour planes of a tile per workgroup: they read overlapping fine neighbourhoods; coarse
// deep texels from the tile's bit masks, so no load waits on the texel type).
// (Measured and not kept: the tile's (4 x 64 + 1) x (4 x 4 + 1) fine window staged in LDS one
// column at a time by row-contiguous loads: 101 instead of 95 us at C2 level 0.)
__device__ __forceinline__ void RestrictDeepTile(int vb, const int* __restrict__ tiles, const unsigned long long* __restrict__ cDeepMask,
                                                 const MixChart* __restrict__ cch, const MixChart* __restrict__ fch, long cplane, long cNd,
                                                 long fplane, long fNd, const float* __restrict__ fres, float* crhs, float* ce, int K) {
  const int row = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTA), tx = threadIdx.x % kTA;
  const unsigned long long* mk = cDeepMask + (long)vb * 4 * kTB + row;
  const int* T = tiles + 3 * vb;
  const MixChart c = cch[T[0]], f = fch[T[0]];
  const int a = T[1] + tx, b = T[2] + row;
  for (int q = 0; q < 4; q++) {
    if (!((mk[q * kTB] >> tx) & 1ull)) continue;  // coarse deep texels from the tile's bit masks
    const int ci = 2 * a + (q & 1), cj = 2 * b + (q >> 1);
    const long idx = (long)q * cplane + c.base2 + (long)b * c.w2 + a;
    const int fi = f.cofx + 2 * (ci - c.cofx), fj = f.cofy + 2 * (cj - c.cofy);
    int n[9];
#pragma unroll
    for (int k = 0; k < 9; k++) n[k] = DenseByte(f, fplane, fi + k % 3 - 1, fj + k / 3 - 1);
    ForColumnChunks(K, [&](int c0, auto kc) {
      constexpr int KC = decltype(kc)::value;
      float v[KC][9];
#pragma unroll
      for (int ch = 0; ch < KC; ch++)
#pragma unroll
        for (int k = 0; k < 9; k++) v[ch][k] = Ld(fres + (c0 + ch) * fNd, n[k]);
#pragma unroll
      for (int ch = 0; ch < KC; ch++) {
        const float* R = v[ch];
        crhs[(long)(c0 + ch) * cNd + idx] = R[4] + (R[1] + R[3] + R[5] + R[7]) * 0.5f + (R[0] + R[2] + R[6] + R[8]) * 0.25f;
        ce[(long)(c0 + ch) * cNd + idx] = 0.0f;
      }
    });
  }
}

// Coarse boundary rhs: the boundary restriction matrix (fine dense columns); kRowLanes lanes per
// coarse row read its entries coalesced, once for up to four columns, summed by shuffles.  An entry
// is one 64-bit word (fine dense column, FP32 weight bits).
__device__ __forceinline__ void RestrictBoundaryRows(int vb, int B, const int* __restrict__ bdIdx, const int* __restrict__ rp,
                                                     const int2* __restrict__ ent, const float* __restrict__ fres, long fNd, float* crhs, float* ce, long cNd,
                                                     int K) {
  const int r = (int)(((long)vb * blockDim.x + threadIdx.x) / kRowLanes), lane = threadIdx.x % kRowLanes;
  if (r >= B) return;  // whole row groups leave together
  const int k0 = rp[r], k1 = rp[r + 1];
  const long d = bdIdx[r];
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = min(4, K - c0);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = k0 + lane; k < k1; k += kRowLanes) {
      const int2 w = ent[k];
      const float v = __int_as_float(w.y);
      const long c = w.x;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q < kc) acc[q] += v * fres[(long)(c0 + q) * fNd + c];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = RowSum(acc[q]);
    if (lane < kc) {
      crhs[(long)(c0 + lane) * cNd + d] = Pick(acc, lane);
      ce[(long)(c0 + lane) * cNd + d] = 0.0f;
    }
  }
}

// Restriction into a level in one launch: blocks [0, nTiles) take the coarse deep tiles, the others
// the coarse boundary rows (disjoint outputs).
__global__ __launch_bounds__(256) void KMixRestrict(int nTiles, const int* __restrict__ tiles, const unsigned long long* __restrict__ cDeepMask,
                                                    const MixChart* __restrict__ cch, const MixChart* __restrict__ fch, long cplane, long cNd,
                                                    long fplane, long fNd, const float* __restrict__ fres, float* crhs, float* ce, int K, int B,
                                                    const int* __restrict__ bdIdx, const int* __restrict__ rp, const int2* __restrict__ ent) {
  if ((int)blockIdx.x < nTiles)
    RestrictDeepTile(XcdTile(blockIdx.x, nTiles), tiles, cDeepMask, cch, fch, cplane, cNd, fplane, fNd, fres, crhs, ce, K);
  else RestrictBoundaryRows(blockIdx.x - nTiles, B, bdIdx, rp, ent, fres, fNd, crhs, ce, cNd, K);
}

// e_fine += P e_coarse at every active fine texel (bilinear, chart-centred coarse grid).  A thread
// owns one fine texel of each of the four colour planes, (2a + qi, 2b + qj); the coarse neighbours of
// all four lie in one 2x2 coarse block (columns ci0, ci0 + 1, rows cj0, cj0 + 1), loaded once where an
// active texel needs it, so every load of a column chunk is in flight together.  Per direction the
// fine texel with even offset from the coarse grid takes one coarse column (base bx), the odd one the
// mean of two (bx, bx + 1): bx = 0 for both when ui0 is even, bx = qi when ui0 is odd.  Deep texels of
// plane `skipQ` are left alone: the Gauss-Seidel sweep that follows overwrites them without reading
// them (skipQ = -1 when no sweep follows).
__global__ __launch_bounds__(256) void KMixProlongAdd(const int* __restrict__ tiles, const unsigned long long* __restrict__ fDeepMask,
                                                      const unsigned long long* __restrict__ fActiveMask, const MixChart* __restrict__ fch,
                                                      const MixChart* __restrict__ cch, long fplane, long fNd, long cplane, long cNd,
                                                      const float* __restrict__ ce, float* fe, int K, int skipQ) {
  // active (and deep) fine texels from the tile's bit masks: no load waits on the texel type
  const int tile = XcdTile(blockIdx.x, gridDim.x), row = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTA), tx = threadIdx.x % kTA;
  const long mo = (long)tile * 4 * kTB + row;
  const int* T = tiles + 3 * tile;
  const MixChart f = fch[T[0]], c = cch[T[0]];
  const int a = T[1] + tx, b = T[2] + row;
  const int ui0 = 2 * a - f.cofx, uj0 = 2 * b - f.cofy;
  const int ci0 = FloorHalf(ui0) + c.cofx, cj0 = FloorHalf(uj0) + c.cofy;
  const bool ei = (ui0 & 1) == 0, ej = (uj0 & 1) == 0;
  int fidx[4];
  bool act[4];
  int need = 0;  // bit cx + 2 cy: coarse block entry (ci0 + cx, cj0 + cy) is read
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int qi = q & 1, qj = q >> 1;
    const long idx = (long)q * fplane + f.base2 + (long)b * f.w2 + a;
    fidx[q] = (int)(4 * idx);
    const bool active = (fActiveMask[mo + q * kTB] >> tx) & 1ull, deep = (fDeepMask[mo + q * kTB] >> tx) & 1ull;
    act[q] = active && !(q == skipQ && deep);
    const int bx = ei ? 0 : qi, by = ej ? 0 : qj;
    const bool ox = ei ? qi == 1 : qi == 0, oy = ej ? qj == 1 : qj == 0;
    if (act[q]) need |= ((1 | (ox ? 2 : 0)) * (1 | (oy ? 4 : 0))) << (bx + 2 * by);
  }
  if (!need) return;
  int nb[4];
#pragma unroll
  for (int n = 0; n < 4; n++) nb[n] = DenseByte(c, cplane, ci0 + (n & 1), cj0 + (n >> 1));
  ForColumnChunks(K, [&](int c0, auto kc) {
    constexpr int KC = decltype(kc)::value;
    float v[KC][4], old[KC][4];
#pragma unroll
    for (int ch = 0; ch < KC; ch++) {
      const float* E = ce + (c0 + ch) * cNd;
      const float* F = fe + (c0 + ch) * fNd;
#pragma unroll
      for (int n = 0; n < 4; n++) v[ch][n] = (need >> n) & 1 ? Ld(E, nb[n]) : 0.0f;
#pragma unroll
      for (int q = 0; q < 4; q++) old[ch][q] = act[q] ? Ld(F, fidx[q]) : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (!act[q]) continue;
      const int qi = q & 1, qj = q >> 1;
      const bool ox = ei ? qi == 1 : qi == 0, oy = ej ? qj == 1 : qj == 0;
      const bool sx = !ei && qi == 1, sy = !ej && qj == 1;  // base column / row is ci0 + 1 / cj0 + 1
#pragma unroll
      for (int ch = 0; ch < KC; ch++) {
        const float* V = v[ch];
        const float b0 = sx ? V[1] : V[0], b1 = sx ? V[3] : V[2];  // base column, rows cj0 and cj0 + 1
        const float v00 = sy ? b1 : b0;
        const float v10 = ox ? (sy ? V[3] : V[1]) : v00;
        const float v01 = oy ? b1 : v00;
        const float v11 = (ox && oy) ? V[3] : (ox ? v10 : v01);
        const float term = (v00 + v10 + v01 + v11) * 0.25f;  // exact scaling: same value as ProlongTerm
        fe[(long)(c0 + ch) * fNd + fidx[q] / 4] = old[ch][q] + term;
      }
    }
  });
}

// Coarsest level: e = A^{-1} rhs with the precomputed dense inverse (FP64), rhs / e in the dense
// FP32 planes; one workgroup per row, which reads the row once for up to four columns (the small
// coarsest level's rhs stays in the L2).
__global__ __launch_bounds__(256) void KMixDenseCoarse(int N, const double* __restrict__ Ainv, const int* __restrict__ g2d,
                                                       const float* __restrict__ rhs, float* e, long Nd, int K) {
  const int i = blockIdx.x;
  const double* row = Ainv + (long)i * N;
  __shared__ double part[4][4];
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = min(4, K - c0);
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
      const double a = row[j];
      const long gj = g2d[j];
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q < kc) s[q] += a * (double)rhs[(long)(c0 + q) * Nd + gj];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      for (int o = 32; o > 0; o >>= 1) s[q] += __shfl_xor(s[q], o, 64);
      if (threadIdx.x % 64 == 0) part[q][threadIdx.x / 64] = s[q];
    }
    __syncthreads();
    if ((int)threadIdx.x < kc) {
      const int q = threadIdx.x;
      e[(long)(c0 + q) * Nd + g2d[i]] = (float)(part[q][0] + part[q][1] + part[q][2] + part[q][3]);
    }
    __syncthreads();
  }
}

__global__ void KMixZero(long n, float* p) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0f;
}

__global__ void KMixZeroD(int n, double* p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// coarsest level: dense FP32 <-> compact FP64
__global__ void KMixGather(int N, const int* __restrict__ g2d, const float* __restrict__ src, long Nd, double* dst, int K) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * K) return;
  const int g = t % N, ch = t / N;
  dst[(long)ch * N + g] = (double)src[(long)ch * Nd + g2d[g]];
}

__global__ void KMixScatter(int N, const int* __restrict__ g2d, const double* __restrict__ src, float* dst, long Nd, int K) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * K) return;
  const int g = t % N, ch = t / N;
  dst[(long)ch * Nd + g2d[g]] = (float)src[(long)ch * N + g];
}

// ------------------------------------------------------------------ FP64 outer loop
static constexpr int kSlots = 64;  // norm accumulators per column (spreads the atomics)

// Block-wide sums of v[0..kc) added to target[c * kSlots + blockIdx.x % kSlots]: one barrier pair for
// all columns of a chunk; every thread of the block must call it with the same kc.
__device__ __forceinline__ void BlockAddSquares4(const double (&v)[4], int kc, double* target) {
  __shared__ double part[4][16];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (c >= kc) break;
    double s = v[c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) part[c][w] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < kc) {
    double s = 0;
    for (int k = 0; k < (int)(blockDim.x / 64); k++) s += part[threadIdx.x][k];
    atomicAdd(target + (long)threadIdx.x * kSlots + (blockIdx.x % kSlots), s);
  }
  __syncthreads();
}

// Block-wide sum of v added to target[blockIdx.x % kSlots]; every thread of the block must call it.
__device__ __forceinline__ void BlockAddSquares(double v, double* target) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __shared__ double part[16];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (lane == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int k = 0; k < (int)(blockDim.x / 64); k++) s += part[k];
    atomicAdd(target + (blockIdx.x % kSlots), s);
  }
  __syncthreads();
}

// Deep texel d's raster line gives the global indices of the rows above and below
// (line = deepLine[d]: start, end, prev, next, first deep index).
__device__ __forceinline__ void DeepRow(const int* __restrict__ deepLine, const int* __restrict__ lines, int d, int& cur, int& prev,
                                        int& next) {
  const int* L = lines + 5 * deepLine[d];
  const int i = d - L[4];
  cur = L[0] + i, prev = L[2] + i, next = L[3] + i;
}

// f(std::integral_constant<int, kc>) for a chunk width kc in 1..kMax known only at run time, so that
// the chunk's loads (every column's) are all issued before the first is used.
template <int kMax = 3, class F>
__device__ __forceinline__ void WithChunkWidth(int kc, F&& f) {
  switch (kc) {
    case 4: if constexpr (kMax >= 4) f(std::integral_constant<int, 4>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    default: f(std::integral_constant<int, 1>()); break;
  }
}

// The nine x values of deep row `cur` (rows prev, cur, next), every column of a chunk.
template <int KC>
__device__ __forceinline__ void RowValues(const double* __restrict__ x, int N, int c0, int prev, int cur, int next, double (&xv)[KC][9]) {
#pragma unroll
  for (int cc = 0; cc < KC; cc++) {
    const double* X = x + (long)(c0 + cc) * N;
#pragma unroll
    for (int k = 0; k < 9; k++) xv[cc][k] = X[(k < 3 ? prev : (k < 6 ? cur : next)) + k % 3 - 1];
  }
}

// The nine couplings of deep row `cur` from the symmetric five-word stencil ([5][N], global indices):
// its own lower couplings and diagonal, and each upper coupling k from slot 8-k of that neighbour.
__device__ __forceinline__ void SymRow(const double* __restrict__ sym, int N, int cur, int next, double cf[9]) {
#pragma unroll
  for (int k = 0; k < 5; k++) cf[k] = sym[(long)k * N + cur];
  cf[5] = sym[3L * N + cur + 1], cf[6] = sym[2L * N + next - 1], cf[7] = sym[(long)N + next], cf[8] = sym[next + 1];
}

// deep rows: r = b - A x with the symmetric five-word FP64 stencil along raster lines; r -> FP32
// dense rhs of the correction, e -> 0, |r|^2 -> norms[ch][slot] (|b|^2 -> bnorms on the first
// residual of a solve).
// Workgroup vb takes kDeepGroups consecutive groups of blockDim deep texels, one texel per thread
// per group, and adds its squared residuals to the norm slots once per column chunk (four times
// fewer atomics than one group per workgroup).  A chunk's loads (every column's) are issued
// together.  (C2, 12.1 M deep rows, 3 columns: 470 us per launch; one group per workgroup 463 us,
// one column per wave 557 us, each row's side values taken from the neighbouring lanes by shuffles
// 561 us, chunks of two columns (92 instead of 116 registers, five waves per SIMD instead of four)
// 520 us, each texel's row indices and dense index from a host-built int4 table instead of its
// raster line (no dependent index loads) 475 against 467 us: the launch time follows neither the
// atomics, the load batching, the load count, the occupancy nor the index chain.)
static constexpr int kDeepGroups = 4;
__device__ __forceinline__ void K64ResidualDeepRows(int vb, int D, const int* __restrict__ deepLine, const int* __restrict__ lines,
                                                    const double* __restrict__ sym, const double* __restrict__ x,
                                                    const double* __restrict__ b, int N, const int* __restrict__ g2d, float* rhs0,
                                                    float* e0, long Nd, double* norms, int K, double* bnorms) {
  for (int c0 = 0; c0 < K; c0 += 3) {  // chunks of three columns: 27 x values in registers
    const int kc = min(3, K - c0);
    double acc[4] = {0, 0, 0, 0}, bacc[4] = {0, 0, 0, 0};
    for (int g = 0; g < kDeepGroups; g++) {
      const int d = (vb * kDeepGroups + g) * blockDim.x + threadIdx.x;
      if (d >= D) break;
      int cur, prev, next;
      DeepRow(deepLine, lines, d, cur, prev, next);
      double cf[9];
      SymRow(sym, N, cur, next, cf);
      const long di = g2d[cur];
      WithChunkWidth<3>(kc, [&](auto kcc) {
        constexpr int KC = decltype(kcc)::value;
        double xv[KC][9], bv[KC];
        RowValues<KC>(x, N, c0, prev, cur, next, xv);
#pragma unroll
        for (int cc = 0; cc < KC; cc++) bv[cc] = b[(long)(c0 + cc) * N + cur];
#pragma unroll
        for (int cc = 0; cc < KC; cc++) {
          const int ch = c0 + cc;
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 9; k++) s += cf[k] * xv[cc][k];
          const double r = bv[cc] - s;
          rhs0[(long)ch * Nd + di] = (float)r;
          e0[(long)ch * Nd + di] = 0.0f;
          acc[cc] += r * r, bacc[cc] += bv[cc] * bv[cc];
        }
      });
    }
    BlockAddSquares4(acc, kc, norms + (long)c0 * kSlots);  // every thread of the block
    if (bnorms) BlockAddSquares4(bacc, kc, bnorms + (long)c0 * kSlots);
  }
}

// The same residual for per-column operators A_c = w[c][0] M + w[c][1] S (parametric solves: one
// refinement loop serves columns with different weights, its FP32 V-cycle built for one nearby pair
// of weights): both level-0 stencils read once per texel for a chunk of columns.
__device__ __forceinline__ void K64ResidualDeepWRows(int vb, int D, const int* __restrict__ deepLine, const int* __restrict__ lines,
                                                     const double* __restrict__ symM, const double* __restrict__ symS,
                                                     const double* __restrict__ x, const double* __restrict__ b, int N,
                                                     const int* __restrict__ g2d, float* rhs0, float* e0, long Nd, double* norms, int K,
                                                     const double* __restrict__ w, double* r64, double* bnorms) {
  for (int c0 = 0; c0 < K; c0 += 2) {  // chunks of two columns (two stencils in registers)
    const int kc = min(2, K - c0);
    double acc[4] = {0, 0, 0, 0}, bacc[4] = {0, 0, 0, 0};
    for (int g = 0; g < kDeepGroups; g++) {
      const int d = (vb * kDeepGroups + g) * blockDim.x + threadIdx.x;
      if (d >= D) break;
      int cur, prev, next;
      DeepRow(deepLine, lines, d, cur, prev, next);
      double cm[9], cs[9];
      SymRow(symM, N, cur, next, cm);
      SymRow(symS, N, cur, next, cs);
      const long di = g2d[cur];
      WithChunkWidth<2>(kc, [&](auto kcc) {
        constexpr int KC = decltype(kcc)::value;
        double xv[KC][9], bv[KC];
        RowValues<KC>(x, N, c0, prev, cur, next, xv);
#pragma unroll
        for (int cc = 0; cc < KC; cc++) bv[cc] = b[(long)(c0 + cc) * N + cur];
#pragma unroll
        for (int cc = 0; cc < KC; cc++) {
          const int ch = c0 + cc;
          double sm = 0.0, ss = 0.0;
#pragma unroll
          for (int k = 0; k < 9; k++) sm += cm[k] * xv[cc][k], ss += cs[k] * xv[cc][k];
          const double r = bv[cc] - (w[2 * ch] * sm + w[2 * ch + 1] * ss);
          if (r64) r64[(long)ch * N + cur] = r;
          rhs0[(long)ch * Nd + di] = (float)r;
          e0[(long)ch * Nd + di] = 0.0f;
          acc[cc] += r * r, bacc[cc] += bv[cc] * bv[cc];
        }
      });
    }
    BlockAddSquares4(acc, kc, norms + (long)c0 * kSlots);  // every thread of the block
    if (bnorms) BlockAddSquares4(bacc, kc, bnorms + (long)c0 * kSlots);
  }
}

// Boundary rows of the per-column-weights residual: rows of M (A_BB, A_BD) then of S.
__device__ __forceinline__ void K64ResidualBoundaryWRows(int vb, int B, const int* __restrict__ bgi, const int* __restrict__ mrpd,
                                                         const int* __restrict__ mcold, const double* __restrict__ mvald,
                                                         const int* __restrict__ mrpb, const int* __restrict__ mcolb,
                                                         const double* __restrict__ mvalb, const int* __restrict__ srpd,
                                                         const int* __restrict__ scold, const double* __restrict__ svald,
                                                         const int* __restrict__ srpb, const int* __restrict__ scolb,
                                                         const double* __restrict__ svalb, const double* __restrict__ x,
                                                         const double* __restrict__ b, int N, const int* __restrict__ g2d, float* rhs0,
                                                         float* e0, long Nd, double* norms, int K, const double* __restrict__ w,
                                                         double* r64, double* bnorms) {
  const int r = (int)(((long)vb * blockDim.x + threadIdx.x) / kRowLanes64), lane = threadIdx.x % kRowLanes64;
  const bool active = r < B;
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = min(4, K - c0);
    double am[4] = {0.0, 0.0, 0.0, 0.0}, as[4] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
      auto rows = [&](const int* rp, const int* cl, const double* vl, bool viaBoundary, double (&acc)[4]) {
        for (int k = rp[r] + lane; k < rp[r + 1]; k += kRowLanes64) {
          const double v = vl[k];
          const int g = viaBoundary ? bgi[cl[k]] : cl[k];
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (q < kc) acc[q] += v * x[(long)(c0 + q) * N + g];
        }
      };
      rows(mrpb, mcolb, mvalb, true, am);
      rows(mrpd, mcold, mvald, false, am);
      rows(srpb, scolb, svalb, true, as);
      rows(srpd, scold, svald, false, as);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) am[q] = RowSum<double, kRowLanes64>(am[q]), as[q] = RowSum<double, kRowLanes64>(as[q]);
    double sq = 0.0, bq = 0.0;
    if (active && lane < kc) {
      const int ch = c0 + lane, g = bgi[r];
      const double bv = b[(long)ch * N + g], rv = bv - (w[2 * ch] * Pick(am, lane) + w[2 * ch + 1] * Pick(as, lane));
      if (r64) r64[(long)ch * N + g] = rv;
      rhs0[(long)ch * Nd + g2d[g]] = (float)rv;
      e0[(long)ch * Nd + g2d[g]] = 0.0f;
      sq = rv * rv, bq = bv * bv;
    }
    {
      const double sqs[4] = {lane == 0 ? sq : 0.0, lane == 1 ? sq : 0.0, lane == 2 ? sq : 0.0, lane == 3 ? sq : 0.0};
      BlockAddSquares4(sqs, kc, norms + (long)c0 * kSlots);
    }
    if (bnorms) {
      const double bqs[4] = {lane == 0 ? bq : 0.0, lane == 1 ? bq : 0.0, lane == 2 ? bq : 0.0, lane == 3 ? bq : 0.0};
      BlockAddSquares4(bqs, kc, bnorms + (long)c0 * kSlots);
    }
  }
}

__device__ __forceinline__ void K64ResidualBoundaryRows(int vb, int B, const int* __restrict__ bgi, const int* __restrict__ rpd,
                                                        const int* __restrict__ cold, const double* __restrict__ vald,
                                                        const int* __restrict__ rpb, const int* __restrict__ colb,
                                                        const double* __restrict__ valb, const double* __restrict__ x,
                                                        const double* __restrict__ b, int N, const int* __restrict__ g2d, float* rhs0,
                                                        float* e0, long Nd, double* norms, int K, double* bnorms) {
  // kRowLanes64 lanes per boundary row (no early exit: every thread takes part in the block sums)
  const int r = (int)(((long)vb * blockDim.x + threadIdx.x) / kRowLanes64), lane = threadIdx.x % kRowLanes64;
  const bool active = r < B;
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = min(4, K - c0);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
      for (int k = rpb[r] + lane; k < rpb[r + 1]; k += kRowLanes64) {
        const double v = valb[k];
        const int g = bgi[colb[k]];
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (q < kc) acc[q] += v * x[(long)(c0 + q) * N + g];
      }
      for (int k = rpd[r] + lane; k < rpd[r + 1]; k += kRowLanes64) {
        const double v = vald[k];
        const int g = cold[k];
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (q < kc) acc[q] += v * x[(long)(c0 + q) * N + g];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = RowSum<double, kRowLanes64>(acc[q]);
    double sq = 0.0, bq = 0.0;
    if (active && lane < kc) {
      const int g = bgi[r];
      const double bv = b[(long)(c0 + lane) * N + g], rv = bv - Pick(acc, lane);
      rhs0[(long)(c0 + lane) * Nd + g2d[g]] = (float)rv;
      e0[(long)(c0 + lane) * Nd + g2d[g]] = 0.0f;
      sq = rv * rv, bq = bv * bv;
    }
    {
      const double sqs[4] = {lane == 0 ? sq : 0.0, lane == 1 ? sq : 0.0, lane == 2 ? sq : 0.0, lane == 3 ? sq : 0.0};
      BlockAddSquares4(sqs, kc, norms + (long)c0 * kSlots);
    }
    if (bnorms) {
      const double bqs[4] = {lane == 0 ? bq : 0.0, lane == 1 ? bq : 0.0, lane == 2 ? bq : 0.0, lane == 3 ? bq : 0.0};
      BlockAddSquares4(bqs, kc, bnorms + (long)c0 * kSlots);
    }
  }
}

// The level-0 FP64 residual in one launch: blocks [0, deepBlocks) take the deep rows, the others
// the boundary rows (disjoint rows; every block takes part in its own norm sums).
__global__ __launch_bounds__(256) void K64Residual(int deepBlocks, int D, const int* __restrict__ deepLine, const int* __restrict__ lines,
                                                   const double* __restrict__ sym, int B, const int* __restrict__ bgi,
                                                   const int* __restrict__ rpd, const int* __restrict__ cold, const double* __restrict__ vald,
                                                   const int* __restrict__ rpb, const int* __restrict__ colb, const double* __restrict__ valb,
                                                   const double* __restrict__ x, const double* __restrict__ b, int N,
                                                   const int* __restrict__ g2d, float* rhs0, float* e0, long Nd, double* norms, int K,
                                                   double* bnorms, int* fresh) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *fresh = 0;  // x is written from here on (read by the next update)
  if ((int)blockIdx.x < deepBlocks)
    K64ResidualDeepRows(XcdTile(blockIdx.x, deepBlocks), D, deepLine, lines, sym, x, b, N, g2d, rhs0, e0, Nd, norms, K, bnorms);
  else
    K64ResidualBoundaryRows(blockIdx.x - deepBlocks, B, bgi, rpd, cold, vald, rpb, colb, valb, x, b, N, g2d, rhs0, e0, Nd, norms, K, bnorms);
}

// The same for per-column weights (M and S rows of both parts).
__global__ __launch_bounds__(256) void K64ResidualW(int deepBlocks, int D, const int* __restrict__ deepLine, const int* __restrict__ lines,
                                                    const double* __restrict__ symM, const double* __restrict__ symS, int B,
                                                    const int* __restrict__ bgi, const int* __restrict__ mrpd, const int* __restrict__ mcold,
                                                    const double* __restrict__ mvald, const int* __restrict__ mrpb,
                                                    const int* __restrict__ mcolb, const double* __restrict__ mvalb,
                                                    const int* __restrict__ srpd, const int* __restrict__ scold,
                                                    const double* __restrict__ svald, const int* __restrict__ srpb,
                                                    const int* __restrict__ scolb, const double* __restrict__ svalb,
                                                    const double* __restrict__ x, const double* __restrict__ b, int N,
                                                    const int* __restrict__ g2d, float* rhs0, float* e0, long Nd, double* norms, int K,
                                                    const double* __restrict__ w, double* r64, double* bnorms) {
  if ((int)blockIdx.x < deepBlocks)
    K64ResidualDeepWRows(XcdTile(blockIdx.x, deepBlocks), D, deepLine, lines, symM, symS, x, b, N, g2d, rhs0, e0, Nd, norms, K, w, r64,
                         bnorms);
  else
    K64ResidualBoundaryWRows(blockIdx.x - deepBlocks, B, bgi, mrpd, mcold, mvald, mrpb, mcolb, mvalb, srpd, scold, svald, srpb, scolb,
                             svalb, x, b, N, g2d, rhs0, e0, Nd, norms, K, w, r64, bnorms);
}

// The first residual of a solve from a zero guess: r = b exactly, so no operator is applied.  r goes
// to the FP32 dense rhs of the correction (e = 0) and, for flexible CG, to r64; |r|^2 = |b|^2 into
// both norm slots.  x is set to zero when writeX is set (flexible CG); refinement leaves x unwritten
// and its first update writes x = 0 + e instead of reading x.  One thread per texel; every thread of
// a block reaches the block sums.
__global__ __launch_bounds__(256) void K64ZeroGuess(int N, const int* __restrict__ g2d, const double* __restrict__ b, double* x, int writeX,
                                                    float* rhs0, float* e0, long Nd, double* norms, double* bnorms, double* r64, int K) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const long di = g < N ? g2d[g] : 0;
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = min(4, K - c0);
    double acc[4] = {0, 0, 0, 0};
    if (g < N)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (q >= kc) break;
        const long o = (long)(c0 + q) * N + g;
        const double v = b[o];
        if (writeX) x[o] = 0.0;
        if (r64) r64[o] = v;
        rhs0[(long)(c0 + q) * Nd + di] = (float)v;
        e0[(long)(c0 + q) * Nd + di] = 0.0f;
        acc[q] = v * v;
      }
    BlockAddSquares4(acc, kc, norms + (long)c0 * kSlots);  // every thread of the block
    if (bnorms) BlockAddSquares4(acc, kc, bnorms + (long)c0 * kSlots);
  }
}

// x += e at every texel: one thread per texel for all columns (its dense index read once).  While
// *fresh is set (the first update of a solve from a zero guess, x not yet written) x = 0 + e, the
// same bits as x += e on a zeroed x, without reading x; the residual launch after it clears *fresh.
__global__ __launch_bounds__(256) void K64Update(int N, const int* __restrict__ g2d, const float* __restrict__ e0, long Nd, double* x, int K,
                                                 const int* __restrict__ fresh) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  const long di = g2d[g];
  if (*fresh) {
#pragma unroll 4
    for (int ch = 0; ch < K; ch++) x[(long)ch * N + g] = 0.0 + (double)e0[(long)ch * Nd + di];
    return;
  }
#pragma unroll 4
  for (int ch = 0; ch < K; ch++) x[(long)ch * N + g] += (double)e0[(long)ch * Nd + di];
}

// ------------------------------------------------------------------ host
MixedSystem::~MixedSystem() {
  for (auto& g : graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  if (stream) (void)hipStreamDestroy(stream);
  if (hostNorms) (void)hipHostFree(hostNorms);
  for (void* p : allocations) (void)hipFree(p);
}

template <class T>
static T* Up(MixedSystem& ms, const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* p = nullptr;
  GDTP_HIP(hipMalloc(&p, v.size() * sizeof(T)));
  ms.allocations.push_back(p);
  ms.bytes += v.size() * sizeof(T);
  GDTP_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

template <class T>
static T* Zeros(MixedSystem& ms, size_t n) {
  T* p = nullptr;
  GDTP_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  ms.allocations.push_back(p);
  ms.bytes += std::max<size_t>(n, 1) * sizeof(T);
  GDTP_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  return p;
}

struct SchwarzBlocks {
  std::vector<int> corePtr{0}, core, extPtr{0}, ext;
  std::vector<long> matPtr{0};
  std::vector<float> mat;
};

// Blocks of up to kBlockCore boundary rows grown breadth-first over the A_BB graph, each extended
// by its A_BB neighbours (the most strongly coupled first, up to kBlockExtMax rows); per block the
// core rows of the inverse of A restricted to the extended rows (FP64 Cholesky, stored FP32).
static void BuildSchwarzBlocks(const Csr& bb, const std::vector<int>& bidx, SchwarzBlocks& out) {
  const int B = bb.rows;
  std::vector<int> owner(B, -1), diagPos(B, -1);
  for (int r = 0; r < B; r++)
    for (int k = bb.rowptr[r]; k < bb.rowptr[r + 1]; k++)
      if (bb.col[k] == r) diagPos[r] = k;
  std::vector<std::vector<int>> cores;
  for (int seed = 0; seed < B; seed++) {
    if (owner[seed] >= 0) continue;
    const int id = (int)cores.size();
    std::vector<int> c{seed};
    owner[seed] = id;
    for (size_t q = 0; q < c.size() && (int)c.size() < kBlockCore; q++)
      for (int k = bb.rowptr[c[q]]; k < bb.rowptr[c[q] + 1] && (int)c.size() < kBlockCore; k++) {
        const int v = bb.col[k];
        if (owner[v] < 0) owner[v] = id, c.push_back(v);
      }
    cores.push_back(std::move(c));
  }
  const int nb = (int)cores.size();
  std::vector<std::vector<int>> exts(nb);
  std::vector<std::vector<float>> mats(nb);
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 4)
  for (int b = 0; b < nb; b++) {
    const std::vector<int>& c = cores[b];
    const int nc = (int)c.size();
    std::vector<std::pair<double, int>> nbr;  // (-strength, row)
    for (int u : c)
      for (int k = bb.rowptr[u]; k < bb.rowptr[u + 1]; k++) {
        const int v = bb.col[k];
        if (owner[v] == b) continue;
        const double st = std::fabs(bb.val[k]) / std::sqrt(bb.val[diagPos[u]] * bb.val[diagPos[v]]);
        nbr.push_back({-st, v});
      }
    std::sort(nbr.begin(), nbr.end());
    std::vector<int>& x = exts[b];
    x = c;
    for (const auto& p : nbr) {
      if ((int)x.size() >= kBlockExtMax) break;
      if (std::find(x.begin() + nc, x.end(), p.second) == x.end()) x.push_back(p.second);
    }
    const int n = (int)x.size();
    std::unordered_map<int, int> pos;
    for (int i = 0; i < n; i++) pos[x[i]] = i;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
      for (int k = bb.rowptr[x[i]]; k < bb.rowptr[x[i] + 1]; k++) {
        auto it = pos.find(bb.col[k]);
        if (it != pos.end()) A[(size_t)i * n + it->second] = bb.val[k];
      }
    // Cholesky A = L L^T (lower triangle in place)
    bool ok = true;
    for (int j = 0; j < n && ok; j++) {
      double d = A[(size_t)j * n + j];
      for (int k = 0; k < j; k++) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
      if (!(d > 0)) { ok = false; break; }
      d = std::sqrt(d);
      A[(size_t)j * n + j] = d;
      for (int i = j + 1; i < n; i++) {
        double v = A[(size_t)i * n + j];
        for (int k = 0; k < j; k++) v -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
        A[(size_t)i * n + j] = v / d;
      }
    }
    if (!ok) {
#pragma omp atomic write
      failed = 1;
      continue;
    }
    std::vector<float>& M = mats[b];
    M.assign((size_t)n * nc, 0.0f);
    std::vector<double> z(n);
    for (int i = 0; i < nc; i++) {  // column i of the inverse (= row i: the inverse is symmetric)
      std::fill(z.begin(), z.end(), 0.0);
      z[i] = 1.0;
      for (int r = 0; r < n; r++) {
        double v = z[r];
        for (int k = 0; k < r; k++) v -= A[(size_t)r * n + k] * z[k];
        z[r] = v / A[(size_t)r * n + r];
      }
      for (int r = n - 1; r >= 0; r--) {
        double v = z[r];
        for (int k = r + 1; k < n; k++) v -= A[(size_t)k * n + r] * z[k];
        z[r] = v / A[(size_t)r * n + r];
      }
      for (int j = 0; j < n; j++) M[(size_t)j * nc + i] = (float)z[j];
    }
  }
  if (failed) throw std::runtime_error("boundary Schwarz block is not positive definite");
  for (int b = 0; b < nb; b++) {
    for (int u : cores[b]) out.core.push_back(bidx[u]);
    out.corePtr.push_back((int)out.core.size());
    out.ext.insert(out.ext.end(), exts[b].begin(), exts[b].end());
    out.extPtr.push_back((int)out.ext.size());
    out.mat.insert(out.mat.end(), mats[b].begin(), mats[b].end());
    out.matPtr.push_back((long)out.mat.size());
  }
}

// The smoother reads its stencil in FP16 only where the median level-0 row sum / diagonal is at
// least this: the FP16 rounding left after the row-sum correction still slows the nearly singular
// geodesic distance system 1e-6 M + S (median ratio ~1e-12) from at most 18 to 28 iterations,
// while screened systems such as M + 1e-3 S at 4096^2 (~3e-5) converge as with FP32.
static constexpr double kHalfStencilRatio = 2e-6;

void UploadMixed(const Hierarchy& h, const OperatorHierarchy& oh, DeviceSystem& ds, int kcap, MixedSystem& ms) {
  const int nl = (int)h.levels.size();
  if (nl < 2) throw std::runtime_error("the mixed-precision solver needs at least two levels");
  ms.levels = nl, ms.kcap = kcap;
  ms.L.assign(nl, DevMixLevel());
  std::vector<std::vector<long>> g2dAll(nl);
  for (int l = 0; l < nl; l++) {
    const Level& H = h.levels[l];
    DevMixLevel& d = ms.L[l];
    d.N = H.numTexels, d.B = H.numBoundary, d.nCharts = (int)H.charts.size();
    d.hcharts.resize(d.nCharts);
    long base = 0;
    for (int c = 0; c < d.nCharts; c++) {
      const GridChart& g = H.charts[c];
      // plane rows padded to 32 floats (128 B): every 64-wide tile row is two whole cache lines
      const int w2 = ((g.width + 1) / 2 + 31) / 32 * 32;
      MixChart m{g.width, g.height, w2, (g.height + 1) / 2, (int)base, g.centerOffset[0], g.centerOffset[1]};
      d.hcharts[c] = m;
      base += (long)m.w2 * m.h2;
    }
    if (base > (1L << 29)) throw std::runtime_error("dense colour plane too large");
    {
      long largest = 0;
      for (const MixChart& m : d.hcharts) largest = std::max(largest, (long)((m.w + 1) / 2) * m.h2);
      d.chartSweeps = largest <= kChartSweepMax;
    }
    d.plane = base, d.Nd = 4 * base;
    if (4 * d.Nd >= (1L << 31)) throw std::runtime_error("texel atlas too large for the mixed-precision solver (2 GiB per dense FP32 column)");
    auto dense = [&](int c, int i, int j) {
      const MixChart& m = d.hcharts[c];
      int q = (i & 1) | ((j & 1) << 1);
      return (long)q * d.plane + m.base2 + (long)(j >> 1) * m.w2 + (i >> 1);
    };
    std::vector<int8_t> type(d.Nd, -1);
    std::vector<int> g2d(d.N);
    std::vector<long>& g2dl = g2dAll[l];
    g2dl.resize(d.N);
    for (int g = 0; g < d.N; g++) {
      const NodeInfo& ni = H.nodeInfo[g];
      long di = dense(ni.chart, ni.ci, ni.cj);
      g2d[g] = (int)di, g2dl[g] = di;
      type[di] = (int8_t)ni.type;
    }
    d.charts = Up(ms, d.hcharts);
    d.g2d = Up(ms, g2d);
    if (l == 0) {  // raster line of every deep texel, for the FP64 outer kernels
      std::vector<int> dl(H.numDeep, -1);
      for (size_t r = 0; r < H.rasterLines.size(); r++) {
        const Line5& x = H.rasterLines[r];
        for (int i = 0; i <= x.end - x.start; i++) dl[x.coeff + i] = (int)r;
      }
      for (int v : dl)
        if (v < 0) throw std::runtime_error("deep texel outside every raster line");
      ms.deepLine = Up(ms, dl);
    }
    d.type = Up(ms, type);
    // tiles of the sub-grid that contain at least one active texel (any plane)
    auto buildTiles = [&](int ta, int tb) {
      std::vector<int> tiles;
      for (int c = 0; c < d.nCharts; c++) {
        const MixChart& m = d.hcharts[c];
        for (int b0 = 0; b0 < m.h2; b0 += tb)
          for (int a0 = 0; a0 < m.w2; a0 += ta) {
            bool any = false;
            for (int q = 0; q < 4 && !any; q++)
              for (int bb = b0; bb < std::min(b0 + tb, m.h2) && !any; bb++)
                for (int aa = a0; aa < std::min(a0 + ta, m.w2) && !any; aa++) {
                  int i = 2 * aa + (q & 1), j = 2 * bb + (q >> 1);
                  if (i < m.w && j < m.h && type[(long)q * d.plane + m.base2 + (long)bb * m.w2 + aa] >= 0) any = true;
                }
            if (any) tiles.insert(tiles.end(), {c, a0, b0});
          }
      }
      return tiles;
    };
    {
      const std::vector<int> tiles = buildTiles(kTA, kTB);
      d.nTiles = (int)tiles.size() / 3;
      d.tiles = Up(ms, tiles);
      // deep and active texels of each tile: [tile][plane][tile row] 64-bit masks, bit = column in the tile
      std::vector<unsigned long long> deep((size_t)d.nTiles * 4 * kTB, 0ull), active(deep.size(), 0ull);
      static_assert(kTA == 64, "one 64-bit mask per tile row");
      for (int t = 0; t < d.nTiles; t++) {
        const MixChart& m = d.hcharts[tiles[3 * t]];
        const int a0 = tiles[3 * t + 1], b0 = tiles[3 * t + 2];
        for (int q = 0; q < 4; q++)
          for (int r = 0; r < kTB && b0 + r < m.h2; r++)
            for (int x = 0; x < kTA && a0 + x < m.w2; x++) {
              const int i = 2 * (a0 + x) + (q & 1), j = 2 * (b0 + r) + (q >> 1);
              const int tp = i < m.w && j < m.h ? type[(long)q * d.plane + m.base2 + (long)(b0 + r) * m.w2 + a0 + x] : -1;
              if (tp == 2) deep[((size_t)t * 4 + q) * kTB + r] |= 1ull << x;
              if (tp >= 0) active[((size_t)t * 4 + q) * kTB + r] |= 1ull << x;
            }
      }
      d.deepMask = Up(ms, deep);
      d.activeMask = Up(ms, active);
    }
    if (l + 1 < nl && !d.chartSweeps) {  // sub-grid rows for the colour-pair sweep: their active spans
      std::vector<int> rows;
      int widest = 0;
      for (int c = 0; c < d.nCharts; c++) {
        const MixChart& m = d.hcharts[c];
        for (int bb = 0; bb < m.h2; bb++) {
          int lo = m.w2, hi = -1;
          for (int q = 0; q < 4; q++)
            for (int aa = 0; aa < m.w2; aa++) {
              int i = 2 * aa + (q & 1), j = 2 * bb + (q >> 1);
              if (i < m.w && j < m.h && type[(long)q * d.plane + m.base2 + (long)bb * m.w2 + aa] >= 0) lo = std::min(lo, aa), hi = std::max(hi, aa);
            }
          if (hi < 0) continue;
          rows.insert(rows.end(), {c, bb, lo, hi + 1});
          widest = std::max(widest, hi + 1 - lo);
        }
      }
      d.pairThreads = widest <= 256 ? 256 : widest <= 512 ? 512 : kPairThreadsMax;
      d.pairSpan = 1;
      while (d.pairSpan <= kPairMaxSpan && d.pairSpan * d.pairThreads < widest) d.pairSpan *= 2;
      if (d.pairSpan > kPairMaxSpan) d.pairSpan = 0;
      if (d.pairSpan == 1) {  // pack whole rows into workgroups, 64 texels per wave
        const int kWaves = d.pairThreads / 64;
        std::vector<int> waves;
        int used = 0;
        for (size_t r = 0; r < rows.size(); r += 4) {
          const int nw = (rows[r + 3] - rows[r + 2] + 63) / 64;
          if (used + nw > kWaves) {
            for (; used < kWaves; used++) waves.insert(waves.end(), {0, 0, 0, 0});
            used = 0;
          }
          for (int k = 0; k < nw; k++) waves.insert(waves.end(), {rows[r], rows[r + 1], rows[r + 2] + 64 * k, rows[r + 3]});
          used += nw;
        }
        for (; used > 0 && used < kWaves; used++) waves.insert(waves.end(), {0, 0, 0, 0});
        rows.swap(waves);
        d.nRows = (int)rows.size() / (4 * kWaves);
        // deep texels of each wave's 64 texels, per plane (bit = lane)
        std::vector<unsigned long long> mask(rows.size(), 0ull);
        for (size_t wv = 0; wv < rows.size() / 4; wv++) {
          const MixChart& m = d.hcharts[rows[4 * wv]];
          const int bb = rows[4 * wv + 1], a0 = rows[4 * wv + 2], a1 = rows[4 * wv + 3];
          for (int q = 0; q < 4; q++)
            for (int x = 0; x < 64 && a0 + x < a1; x++) {
              const int i = 2 * (a0 + x) + (q & 1), j = 2 * bb + (q >> 1);
              if (i < m.w && j < m.h && type[(long)q * d.plane + m.base2 + (long)bb * m.w2 + a0 + x] == 2) mask[4 * wv + q] |= 1ull << x;
            }
        }
        d.rowMask = Up(ms, mask);
      } else {
        d.nRows = (int)rows.size() / 4;
      }
      d.rows = Up(ms, rows);
    }
    if (l + 1 < nl && !d.chartSweeps)
      for (int sh = 0; sh < kFusedShapes; sh++) {
        int same = -1;  // shapes of equal tile size share one tile list
        for (int o = 0; o < sh && same < 0; o++)
          if (kFusedShape[o].ta == kFusedShape[sh].ta && kFusedShape[o].tb == kFusedShape[sh].tb) same = o;
        if (same >= 0) {
          d.nFTiles[sh] = d.nFTiles[same], d.ftiles[sh] = d.ftiles[same];
          continue;
        }
        const std::vector<int> tiles = buildTiles(kFusedShape[sh].ta, kFusedShape[sh].tb);
        d.nFTiles[sh] = (int)tiles.size() / 3;
        d.ftiles[sh] = Up(ms, tiles);
      }
    // deep stencils (reciprocal form) on smoothed levels
    if (l + 1 < nl) {
      std::vector<float> st(10 * (size_t)d.Nd, 0.0f);
      const std::vector<double>& rc = oh.levels[l].deepRecip;
      for (int dd = 0; dd < H.numDeep; dd++) {
        long di = g2dl[H.deepGlobalIndex[dd]];
        for (int k = 0; k < 10; k++) st[(size_t)k * d.Nd + di] = (float)rc[10 * (size_t)dd + k];
      }
      if (l == 0 && H.numDeep > 0) {
        std::vector<double> ratio(H.numDeep);
        for (int dd = 0; dd < H.numDeep; dd++) {
          double s = 1.0;
          for (int k = 0; k < 9; k++)
            if (k != 4) s += rc[10 * (size_t)dd + k];
          ratio[dd] = s;
        }
        std::nth_element(ratio.begin(), ratio.begin() + ratio.size() / 2, ratio.end());
        ms.rowSumRatio = ratio[ratio.size() / 2];
      }
      const bool half = ms.rowSumRatio >= kHalfStencilRatio;
      if (!half) d.stencil = Up(ms, st);
      // residuals: raw A in symmetric five-word form.  A deep texel reads its upper coupling k = 5..8
      // from slot 8-k of that neighbour; a neighbour that is not deep gets that slot from the deep
      // texel's own stencil (one writer per slot).
      {
        const std::vector<double>& raw = oh.levels[l].deepRaw;
        std::vector<float> sy(5 * (size_t)d.Nd, 0.0f);
        bool inside = true;
#pragma omp parallel for schedule(static) reduction(&& : inside)
        for (int dd = 0; dd < H.numDeep; dd++) {
          const NodeInfo& ni = H.nodeInfo[H.deepGlobalIndex[dd]];
          const MixChart& m = d.hcharts[ni.chart];
          if (ni.ci < 1 || ni.cj < 1 || ni.ci + 1 >= m.w || ni.cj + 1 >= m.h) {
            inside = false;
            continue;
          }
          const long di = dense(ni.chart, ni.ci, ni.cj);
          for (int k = 0; k < 5; k++) sy[k * (size_t)d.Nd + di] = (float)raw[10 * (size_t)dd + k];
          for (int k = 5; k < 9; k++) {
            const long qn = dense(ni.chart, ni.ci + k % 3 - 1, ni.cj + k / 3 - 1);
            if (type[qn] != 2) sy[(8 - k) * (size_t)d.Nd + qn] = (float)raw[10 * (size_t)dd + k];
          }
        }
        if (!inside) throw std::runtime_error("deep texel on the edge of its chart grid");
        d.stencilSym = Up(ms, sy);
      }
      // The smoother's off-diagonal ratios in FP16 (|c_k| <= ~1, relative rounding 2^-11) and, in
      // FP32, c4 and the rounding sum s = sum_k (c_k - fp16(c_k)).  The sweep subtracts s times the
      // mean of the eight neighbours, which restores the row sum: the FP16 rounding then only acts
      // on the local variation of e, not on its smooth part (a row-sum error would inject noise of
      // 1e-4 |e| per sweep, fatal on nearly singular operators such as 1e-6 M + S).  Residuals keep
      // the FP32 stencil above.
      if (half) {
        std::vector<uint16_t> sh(8 * (size_t)d.Nd, 0);
        std::vector<float> sc(2 * (size_t)d.Nd, 0.0f);
        bool range = true;
        // largest |ratio| kept in FP16: the format's range; GDTP_FP16_RATIO_LIMIT lowers it (tests
        // use it to force the FP32 fallback on ordinary operators)
        float limit = 65504.0f;
        if (const char* env = std::getenv("GDTP_FP16_RATIO_LIMIT")) limit = std::min(limit, (float)std::atof(env));
#pragma omp parallel for schedule(static) reduction(&& : range)
        for (long t = 0; t < d.Nd; t++) {
          double round = 0.0;
          for (int k = 0, o = 0; k < 9; k++) {
            if (k == 4) continue;
            const float v = st[(size_t)k * d.Nd + t];
            if (!(std::fabs(v) < limit)) range = false;
            const _Float16 h = (_Float16)v;
            sh[8 * t + o++] = __builtin_bit_cast(uint16_t, h);
            round += (double)v - (double)(float)h;
          }
          sc[2 * t] = st[4 * (size_t)d.Nd + t], sc[2 * t + 1] = (float)round;
        }
        if (range) {
          d.stencilH = reinterpret_cast<uint4*>(Up(ms, sh));
          d.stencilC = reinterpret_cast<float2*>(Up(ms, sc));
        } else {  // a ratio beyond the FP16 range: this level keeps the FP32 smoother
          d.stencil = Up(ms, st);
        }
      }
      // boundary rows: merged A_BB (off-diagonal) and A_BD with dense columns
      const Csr &bb = oh.levels[l].bb, &bd = oh.levels[l].bd;
      std::vector<int> rp(1, 0), cl, bidx(d.B);
      std::vector<float> vl, dg(d.B, 0.0f);
      for (int r = 0; r < d.B; r++) {
        bidx[r] = g2d[H.boundaryGlobalIndex[r]];
        for (int k = bb.rowptr[r]; k < bb.rowptr[r + 1]; k++) {
          if (bb.col[k] == r) { dg[r] = (float)bb.val[k]; continue; }
          cl.push_back(g2d[H.boundaryGlobalIndex[bb.col[k]]]);
          vl.push_back((float)bb.val[k]);
        }
        for (int k = bd.rowptr[r]; k < bd.rowptr[r + 1]; k++) cl.push_back(g2d[bd.col[k]]), vl.push_back((float)bd.val[k]);
        rp.push_back((int)cl.size());
        if (!(dg[r] > 0.0f)) throw std::runtime_error("boundary row without positive diagonal");
      }
      std::vector<int2> ent(cl.size());
      for (size_t k = 0; k < cl.size(); k++) ent[k] = make_int2(cl[k], __builtin_bit_cast(int, vl[k]));
      d.bdIdx = Up(ms, bidx), d.bRowptr = Up(ms, rp), d.bEnt = Up(ms, ent), d.bDiag = Up(ms, dg);
      SchwarzBlocks blk;
      BuildSchwarzBlocks(bb, bidx, blk);
      d.nBlocks = (int)blk.corePtr.size() - 1;
      d.blkCorePtr = Up(ms, blk.corePtr), d.blkCore = Up(ms, blk.core), d.blkExtPtr = Up(ms, blk.extPtr), d.blkExt = Up(ms, blk.ext);
      d.blkMatPtr = Up(ms, blk.matPtr);
      // FP16 quads of extended rows, each core row scaled by its largest |entry|, where the step
      // streams its matrices; levels with at most one block per CU prefetch every entry and are
      // latency-bound, and keep FP32 (FP16 pairs' conversions made them 0.5-0.8 us slower).
      // GDTP_SCHWARZ_HALF_BLOCKS lowers the block count from which levels take the quads (tests use
      // it to run small systems through them).
      int halfBlocks = kSchwarzPrefetchBlocks + 1;
      if (const char* env = std::getenv("GDTP_SCHWARZ_HALF_BLOCKS")) halfBlocks = std::atoi(env);
      d.blkMat = Up(ms, blk.mat);
      if (d.stencilH && d.nBlocks >= halfBlocks) {
        std::vector<long> hp(1, 0);
        for (int b = 0; b < d.nBlocks; b++)
          hp.push_back(hp.back() + (long)((blk.extPtr[b + 1] - blk.extPtr[b] + 3) / 4) * (blk.corePtr[b + 1] - blk.corePtr[b]));
        std::vector<uint2> mh(hp.back());
        std::vector<float> sc(blk.core.size(), 1.0f);
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = 0; b < d.nBlocks; b++) {
          const int c0 = blk.corePtr[b], nc = blk.corePtr[b + 1] - c0, nx = blk.extPtr[b + 1] - blk.extPtr[b];
          const float* M = blk.mat.data() + blk.matPtr[b];
          uint2* H = mh.data() + hp[b];
          for (int i = 0; i < nc; i++) {
            float mx = 0.0f;
            for (int j = 0; j < nx; j++) mx = std::max(mx, std::fabs(M[(size_t)j * nc + i]));
            const float sr = mx > 0.0f ? mx : 1.0f;
            sc[c0 + i] = sr;
            auto h = [&](int j) { return j < nx ? (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(M[(size_t)j * nc + i] / sr)) : 0u; };
            for (int sl = 0; 4 * sl < nx; sl++)
              H[(size_t)sl * nc + i] = make_uint2(h(4 * sl) | h(4 * sl + 1) << 16, h(4 * sl + 2) | h(4 * sl + 3) << 16);
          }
        }
        d.blkMatPtrH = Up(ms, hp), d.blkMatH = Up(ms, mh), d.blkScale = Up(ms, sc);
      }
    } else {
      std::vector<int> bidx(d.B);
      for (int r = 0; r < d.B; r++) bidx[r] = g2d[H.boundaryGlobalIndex[r]];
      d.bdIdx = Up(ms, bidx);
    }
    // restriction into this level (rows: this level's boundary texels; columns: fine dense)
    if (l > 0) {
      const Csr& R = h.levels[l - 1].boundaryRestriction;
      std::vector<int2> ent(R.col.size());
      for (size_t k = 0; k < R.col.size(); k++) ent[k] = make_int2((int)g2dAll[l - 1][R.col[k]], __builtin_bit_cast(int, (float)R.val[k]));
      d.rRowptr = Up(ms, R.rowptr), d.rEnt = Up(ms, ent);
    }
    // every active texel's bilinear parents must be active texels of the next level
    if (l + 1 < nl) {
      const Level& C = h.levels[l + 1];
      for (int g = 0; g < d.N; g++) {
        const NodeInfo& ni = H.nodeInfo[g];
        const GridChart &fc = H.charts[ni.chart], &cc = C.charts[ni.chart];
        int ui = ni.ci - fc.centerOffset[0], uj = ni.cj - fc.centerOffset[1];
        int ci = (ui >= 0 ? ui / 2 : -((1 - ui) / 2)) + cc.centerOffset[0], cj = (uj >= 0 ? uj / 2 : -((1 - uj) / 2)) + cc.centerOffset[1];
        for (int dj = 0; dj <= (uj & 1); dj++)
          for (int di = 0; di <= (ui & 1); di++) {
            int a = ci + di, bq = cj + dj;
            if (a < 0 || bq < 0 || a >= cc.width || bq >= cc.height || cc.node(a, bq) < 0)
              throw std::runtime_error("prolongation parent outside the coarse chart");
          }
      }
    }
    d.e = d.eHome = Zeros<float>(ms, (size_t)d.Nd * kcap);
    if (d.ftiles[0]) d.e2 = Zeros<float>(ms, (size_t)d.Nd * kcap);
    d.rhs = Zeros<float>(ms, (size_t)d.Nd * kcap);
    d.res = Zeros<float>(ms, (size_t)d.Nd * kcap);
    d.tmpB = Zeros<float>(ms, (size_t)std::max(d.B, 1) * kcap);
  }
  const size_t before = UploadByteCounter();
  UploadLdlt(ms.allocations, oh.coarseFactor, ms.coarse, kcap);
  ms.bytes += UploadByteCounter() - before;
  const int Nc = ms.L[nl - 1].N;
  if (Nc <= kDenseCoarseMax) {  // dense inverse of the coarsest operator, by columns of the LDL^T solve
    std::vector<double> inv((size_t)Nc * Nc);
#pragma omp parallel
    {
      std::vector<double> unit((size_t)Nc, 0.0), col((size_t)Nc);
#pragma omp for schedule(dynamic, 16)
      for (int j = 0; j < Nc; j++) {
        unit[j] = 1.0;
        SolveLdlt(oh.coarseFactor, unit.data(), col.data(), 1, Nc);
        unit[j] = 0.0;
        for (int i = 0; i < Nc; i++) inv[(size_t)i * Nc + j] = col[i];
      }
    }
    ms.coarseInv = Up(ms, inv);
  }
  ms.coarseRhs = Zeros<double>(ms, (size_t)Nc * kcap);
  ms.coarseX = Zeros<double>(ms, (size_t)Nc * kcap);
  ms.coarseY = Zeros<double>(ms, (size_t)Nc * kcap);
  ms.norms = Zeros<double>(ms, 2 * (size_t)kcap * kSlots);
  ms.weights = Zeros<double>(ms, 2 * (size_t)kcap);
  {
    const size_t n0 = (size_t)ms.L[0].N * kcap;
    ms.r64 = Zeros<double>(ms, n0), ms.p64 = Zeros<double>(ms, n0), ms.q64 = Zeros<double>(ms, n0), ms.z64 = Zeros<double>(ms, n0);
    ms.dots = Zeros<double>(ms, 3 * (size_t)kcap * kDotBlocks);
    ms.alpha = Zeros<double>(ms, kcap), ms.beta = Zeros<double>(ms, kcap), ms.pqPrev = Zeros<double>(ms, kcap);
    ms.first = Zeros<int>(ms, 1);
    ms.fresh = Zeros<int>(ms, 1);
  }
  GDTP_HIP(hipHostMalloc((void**)&ms.hostNorms, 2 * (size_t)kcap * kSlots * sizeof(double), hipHostMallocDefault));
  GDTP_HIP(hipStreamCreateWithFlags(&ms.stream, hipStreamNonBlocking));
  (void)ds;
}

// One colour of the deep sweep: FP16 stencil where the level holds one (well-screened operators),
// else FP32.
static void DeepGsLaunch(DevMixLevel& d, int q, int K, hipStream_t s) {
  if (d.stencilH)
    KMixDeepGs<true><<<d.nTiles, 256, 0, s>>>(q, d.nTiles, d.tiles, d.deepMask, d.charts, d.plane, d.Nd, d.stencil, d.stencilC, d.stencilH, d.e, d.rhs, K);
  else
    KMixDeepGs<false><<<d.nTiles, 256, 0, s>>>(q, d.nTiles, d.tiles, d.deepMask, d.charts, d.plane, d.Nd, d.stencil, d.stencilC, d.stencilH, d.e, d.rhs, K);
}

// One colour pair (first colour qa) of the deep sweep over the level's rows, columns in chunks of up
// to four.
template <bool kHalf>
static void PairLaunch(DevMixLevel& d, int qa, int K, hipStream_t s) {
  for (int c0 = 0; c0 < K; c0 += 4) {
    const long off = (long)c0 * d.Nd;
#define GDTP_PAIR(KC, NT, P)                                                                                                \
  KMixPairGs<kHalf, KC, NT, P><<<d.nRows, NT, 0, s>>>(qa, d.rows, d.rowMask, d.nRows, d.charts, d.plane, d.Nd, d.stencil,      \
                                                      d.stencilC, d.stencilH, d.e + off, d.rhs + off)
#define GDTP_PAIR_SPAN(KC)                                                                \
  if (d.pairSpan == 1 && d.pairThreads == 256) GDTP_PAIR(KC, 256, 1);                    \
  else if (d.pairSpan == 1 && d.pairThreads == 512) GDTP_PAIR(KC, 512, 1);               \
  else if (d.pairSpan == 1) GDTP_PAIR(KC, kPairThreadsMax, 1);                          \
  else GDTP_PAIR(KC, kPairThreadsMax, kPairMaxSpan);
    switch (std::min(4, K - c0)) {
      case 4: GDTP_PAIR_SPAN(4); break;
      case 3: GDTP_PAIR_SPAN(3); break;
      case 2: GDTP_PAIR_SPAN(2); break;
      default: GDTP_PAIR_SPAN(1); break;
    }
#undef GDTP_PAIR_SPAN
#undef GDTP_PAIR
  }
}

// One fused four-colour sweep (KMixDeepGsFused) over the level's tiles of shape `shape`, columns in
// chunks of up to four; e and e2 swap afterwards.
template <bool kHalf, bool kPost, bool kProlong, int I>
static void FusedLaunch(DevMixLevel& d, int K, hipStream_t s, FusedProlong pr) {
  constexpr FusedShapeInfo F = kFusedShape[I];
  for (int c0 = 0; c0 < K; c0 += 4) {
    const long off = (long)c0 * d.Nd;
    FusedProlong p = pr;
    if (p.ce) p.ce += (long)c0 * p.cNd;
#define GDTP_FUSED(KC)                                                                                                               \
  KMixDeepGsFused<kHalf, kPost, kProlong, F.ta, F.tb, F.threads, KC, F.variant & 3, (F.variant & kFusedXcd) != 0>                  \
      <<<d.nFTiles[I], F.threads, 0, s>>>(d.ftiles[I], d.charts, d.plane, d.Nd, d.stencil, d.stencilC, d.stencilH, d.type, d.e + off, \
                                          d.e2 + off, d.rhs + off, p)
    switch (std::min(4, K - c0)) {
      case 4: GDTP_FUSED(4); break;
      case 3: GDTP_FUSED(3); break;
      case 2: GDTP_FUSED(2); break;
      default: GDTP_FUSED(1); break;
    }
#undef GDTP_FUSED
  }
  std::swap(d.e, d.e2);
}

template <bool kHalf, bool kPost, bool kProlong>
static void FusedShapeLaunch(DevMixLevel& d, int shape, int K, hipStream_t s, FusedProlong pr) {
  static_assert(kFusedShapes == 7, "one case per fused tile shape");
#define GDTP_SHAPE(I) FusedLaunch<kHalf, kPost, kProlong, I>(d, K, s, pr)
  switch (shape) {
    case 0: GDTP_SHAPE(0); break;
    case 1: GDTP_SHAPE(1); break;
    case 2: GDTP_SHAPE(2); break;
    case 3: GDTP_SHAPE(3); break;
    case 4: GDTP_SHAPE(4); break;
    case 5: GDTP_SHAPE(5); break;
    default: GDTP_SHAPE(6); break;
  }
#undef GDTP_SHAPE
}

// Whether the fused sweep of level d runs (a fused tile shape with tiles on this level).
static bool Fused(const DevMixLevel& d) { return d.shape >= 0 && !d.chartSweeps && d.nFTiles[d.shape] > 0; }

// One deep four-colour sweep of a tiled level (colours 0..3, post: 3..0): fused when the system
// uses a fused tile shape, one launch per colour pair with kShapePairs (where the widest row fits),
// else one launch per colour.  With `prolong` (post, fused only) the
// sweep first adds the prolongation of the coarse correction.  Returns the launches made.
static int DeepSweep(MixedSystem& ms, DevMixLevel& d, bool post, int K, hipStream_t s, const FusedProlong* prolong = nullptr) {
  const int shape = d.shape;
  if (Fused(d)) {
    const FusedProlong pr = prolong ? *prolong : FusedProlong();
    if (prolong) {
      if (d.stencilH) FusedShapeLaunch<true, true, true>(d, shape, K, s, pr);
      else FusedShapeLaunch<false, true, true>(d, shape, K, s, pr);
    } else if (d.stencilH) {
      post ? FusedShapeLaunch<true, true, false>(d, shape, K, s, pr) : FusedShapeLaunch<true, false, false>(d, shape, K, s, pr);
    } else {
      post ? FusedShapeLaunch<false, true, false>(d, shape, K, s, pr) : FusedShapeLaunch<false, false, false>(d, shape, K, s, pr);
    }
    return (K + 3) / 4;
  }
  if (d.shape == kShapePairs && d.pairSpan > 0 && d.nRows > 0) {  // pairs (0, 1), (2, 3); post (3, 2), (1, 0)
    for (const int qa : {post ? 3 : 0, post ? 1 : 2}) d.stencilH ? PairLaunch<true>(d, qa, K, s) : PairLaunch<false>(d, qa, K, s);
    return 2 * ((K + 3) / 4);
  }
  for (int qq = 0; qq < 4; qq++) DeepGsLaunch(d, post ? 3 - qq : qq, K, s);
  return 4;
}

// res: the boundary rows' residual, or null when e is zero on the level (the residual is the rhs)
template <int W, int kPre>
static void SchwarzLaunch(MixedSystem& ms, DevMixLevel& d, const float* res, int K, hipStream_t s) {
  if (d.blkMatH && ms.schwarzHalf)
    KMixSchwarz<W, kPre / 4, uint2><<<d.nBlocks, 64 * W, 0, s>>>(d.blkCorePtr, d.blkCore, d.blkExtPtr, d.blkExt, d.blkMatPtrH, d.blkMatH,
                                                                d.blkScale, res, d.bdIdx, d.rhs, d.B, d.e, d.Nd, K, ms.omega);
  else
    KMixSchwarz<W, kPre, float><<<d.nBlocks, 64 * W, 0, s>>>(d.blkCorePtr, d.blkCore, d.blkExtPtr, d.blkExt, d.blkMatPtr, d.blkMat,
                                                            nullptr, res, d.bdIdx, d.rhs, d.B, d.e, d.Nd, K, ms.omega);
}

// zeroE: e is zero on the whole level (start of pre-smoothing), so the first sweep takes the rhs
// as its residual without a residual pass.
static void BoundarySweeps(MixedSystem& ms, DevMixLevel& d, int K, hipStream_t s, bool zeroE) {
  if (d.B == 0 || d.nBlocks == 0) return;
  for (int sw = 0; sw < ms.boundarySweeps; sw++) {
    const bool fromRhs = zeroE && sw == 0;
    if (!fromRhs)
      KMixBoundaryResidual<<<NBlk((long)d.B * kRowLanes, 256), 256, 0, s>>>(d.B, d.bdIdx, d.bRowptr, d.bEnt, d.bDiag, d.e, d.rhs,
                                                                            d.tmpB, 1, d.Nd, K);
    const float* res = fromRhs ? nullptr : d.tmpB;
    if (d.nBlocks <= kSchwarzPrefetchBlocks)
      SchwarzLaunch<kSchwarzWaves, kBlockExtMax / kSchwarzWaves>(ms, d, res, K, s);
    else if (d.nBlocks < kSchwarzManyBlocks)
      SchwarzLaunch<kSchwarzWaves, 0>(ms, d, res, K, s);
    else
      SchwarzLaunch<kSchwarzWavesLarge, 0>(ms, d, res, K, s);
  }
}

// Pre-smoothing: (boundary Schwarz, deep colours 0..3) x nuPre; post-smoothing: (deep colours
// 3..0) x nuPost, then one boundary Schwarz step (a Schwarz step after every post sweep converged
// marginally faster per V-cycle but cost more per digit: C2 15.1 against 13.9 ms per solve).
// zeroE: e is zero on the whole level when pre-smoothing starts.
// prolong (post-smoothing of a fused level): the first sweep adds the coarse correction itself.
static void Smooth(MixedSystem& ms, int l, int K, bool post, hipStream_t s, bool zeroE = false, const FusedProlong* prolong = nullptr) {
  DevMixLevel& d = ms.L[l];
  const int nu = post ? ms.nuPost : ms.nuPre;
  for (int it = 0; it < nu; it++) {
    if (!post) BoundarySweeps(ms, d, K, s, zeroE && it == 0);
    if (d.nTiles > 0 && d.chartSweeps) {
      // post-smoothing's deep sweeps follow one another with nothing between them: one launch
      const int sweeps = post ? nu : 1;
      if (d.stencilH)
        KMixChartGs<true><<<d.nCharts, 1024, 0, s>>>(post ? 1 : 0, sweeps, d.charts, d.plane, d.Nd, d.stencil, d.stencilC, d.stencilH, d.e, d.rhs, K);
      else
        KMixChartGs<false><<<d.nCharts, 1024, 0, s>>>(post ? 1 : 0, sweeps, d.charts, d.plane, d.Nd, d.stencil, d.stencilC, d.stencilH, d.e, d.rhs, K);
      if (post) it = nu - 1;
    } else if (d.nTiles > 0) {
      DeepSweep(ms, d, post, K, s, it == 0 ? prolong : nullptr);
    }
    if (post && it == nu - 1) BoundarySweeps(ms, d, K, s, false);
  }
}

void InvalidateGraph(MixedSystem& ms) {
  for (auto& g : ms.graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  ms.graphs.clear();
}

static constexpr size_t kMaxGraphs = 8;
// automatic V-cycles per iteration: 2 when the level-0 row-sum ratio is at least this, else 1
static constexpr double kTwoCycleRatio = 2e-6;

int MixedDeepSweep(MixedSystem& ms, int l, int K, hipStream_t s) {
  if (l < 0 || l + 1 >= ms.levels) throw std::runtime_error("no smoother on this level");
  DevMixLevel& d = ms.L[l];
  if (d.chartSweeps) throw std::runtime_error("level swept one chart per workgroup, not by tiles");
  return DeepSweep(ms, d, false, K, s);
}

// zeroE0: the level-0 correction is zero on entry (first V-cycle after a residual was written);
// coarser levels always start from zero (the restriction zeroes e where it writes the rhs).
void MixedVCycle(DeviceSystem& ds, MixedSystem& ms, int K, hipStream_t s, bool zeroE0) {
  const int nl = ms.levels;
  for (int l = 0; l + 1 < nl; l++) {
    DevMixLevel &d = ms.L[l], &c = ms.L[l + 1];
    Smooth(ms, l, K, false, s, l > 0 || zeroE0);
    const long rb = d.B > 0 ? NBlk((long)d.B * kRowLanes, 256) : 0, cb = c.B > 0 ? NBlk((long)c.B * kRowLanes, 256) : 0;
    if (d.nTiles + rb > 0)
      (d.nTiles > kResidualWideTiles ? KMixResidual<2> : KMixResidual<4>)<<<d.nTiles + rb, 256, 0, s>>>(
          d.nTiles, d.tiles, d.deepMask, d.charts, d.plane, d.Nd, d.stencilSym, d.e, d.rhs, d.res, K, ms.nuPre > 0 && !d.stencilH ? 3 : -1,
          d.B, d.bdIdx, d.bRowptr, d.bEnt, d.bDiag);
    if (c.nTiles + cb > 0)
      KMixRestrict<<<c.nTiles + cb, 256, 0, s>>>(c.nTiles, c.tiles, c.deepMask, c.charts, d.charts, c.plane, c.Nd, d.plane, d.Nd, d.res, c.rhs,
                                                 c.e, K, c.B, c.bdIdx, c.rRowptr, c.rEnt);
  }
  {
    DevMixLevel& c = ms.L[nl - 1];
    const long nk = (long)c.N * K;
    if (ms.coarseInv) {
      KMixDenseCoarse<<<c.N, 256, 0, s>>>(c.N, ms.coarseInv, c.g2d, c.rhs, c.e, c.Nd, K);
    } else {
      KMixGather<<<NBlk(nk, 256), 256, 0, s>>>(c.N, c.g2d, c.rhs, c.Nd, ms.coarseRhs, K);
      LdltSolve(ms.coarse, ms.coarseRhs, c.N, ms.coarseX, c.N, nullptr, ms.coarseY, K, s);
      KMixScatter<<<NBlk(nk, 256), 256, 0, s>>>(c.N, c.g2d, ms.coarseX, c.e, c.Nd, K);
    }
  }
  for (int l = nl - 2; l >= 0; l--) {
    DevMixLevel &d = ms.L[l], &c = ms.L[l + 1];
    // the first post-smoothing sweep of a fused level adds the prolongation itself
    const bool fuse = ms.fuseProlong && ms.nuPost > 0 && Fused(d);
    if (d.nTiles > 0 && !fuse)
      KMixProlongAdd<<<d.nTiles, 256, 0, s>>>(d.tiles, d.deepMask, d.activeMask, d.charts, c.charts, d.plane, d.Nd, c.plane, c.Nd, c.e, d.e,
                                              K, ms.nuPost > 0 ? 3 : -1);
    FusedProlong pr;
    pr.cch = c.charts, pr.cplane = c.plane, pr.cNd = c.Nd, pr.ce = c.e;
    Smooth(ms, l, K, true, s, false, fuse ? &pr : nullptr);
  }
  (void)ds;
}

// ------------------------------------------------------------------ flexible conjugate gradients
// q = A p for per-column operators A_c = w[c][0] M + w[c][1] S (level 0; deep rows by raster line)
__device__ __forceinline__ void K64ApplyDeepWRows(int vb, int D, const int* __restrict__ deepLine, const int* __restrict__ lines,
                                                  const double* __restrict__ symM, const double* __restrict__ symS,
                                                  const double* __restrict__ p, int N, double* q, int K, const double* __restrict__ w) {
  const int d = vb * blockDim.x + threadIdx.x;
  if (d >= D) return;
  int cur, prev, next;
  DeepRow(deepLine, lines, d, cur, prev, next);
  double cm[9], cs[9];
  SymRow(symM, N, cur, next, cm);
  SymRow(symS, N, cur, next, cs);
  for (int c0 = 0; c0 < K; c0 += 3)
    WithChunkWidth<3>(min(3, K - c0), [&](auto kcc) {
      constexpr int KC = decltype(kcc)::value;
      double pv[KC][9];
      RowValues<KC>(p, N, c0, prev, cur, next, pv);
#pragma unroll
      for (int cc = 0; cc < KC; cc++) {
        const int ch = c0 + cc;
        double sm = 0.0, ss = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++) sm += cm[k] * pv[cc][k], ss += cs[k] * pv[cc][k];
        q[(long)ch * N + cur] = w[2 * ch] * sm + w[2 * ch + 1] * ss;
      }
    });
}

__device__ __forceinline__ void K64ApplyBoundaryWRows(int vb, int B, const int* __restrict__ bgi, const int* __restrict__ mrpd,
                                                      const int* __restrict__ mcold, const double* __restrict__ mvald,
                                                      const int* __restrict__ mrpb, const int* __restrict__ mcolb,
                                                      const double* __restrict__ mvalb, const int* __restrict__ srpd,
                                                      const int* __restrict__ scold, const double* __restrict__ svald,
                                                      const int* __restrict__ srpb, const int* __restrict__ scolb,
                                                      const double* __restrict__ svalb, const double* __restrict__ p, int N, double* q,
                                                      int K, const double* __restrict__ w) {
  const int r = (int)(((long)vb * blockDim.x + threadIdx.x) / kRowLanes64), lane = threadIdx.x % kRowLanes64;
  if (r >= B) return;  // whole row groups leave together
  for (int c0 = 0; c0 < K; c0 += 4) {
    const int kc = min(4, K - c0);
    double am[4] = {0.0, 0.0, 0.0, 0.0}, as[4] = {0.0, 0.0, 0.0, 0.0};
    auto rows = [&](const int* rp, const int* cl, const double* vl, bool viaBoundary, double (&acc)[4]) {
      for (int k = rp[r] + lane; k < rp[r + 1]; k += kRowLanes64) {
        const double v = vl[k];
        const int g = viaBoundary ? bgi[cl[k]] : cl[k];
#pragma unroll
        for (int qq = 0; qq < 4; qq++)
          if (qq < kc) acc[qq] += v * p[(long)(c0 + qq) * N + g];
      }
    };
    rows(mrpb, mcolb, mvalb, true, am);
    rows(mrpd, mcold, mvald, false, am);
    rows(srpb, scolb, svalb, true, as);
    rows(srpd, scold, svald, false, as);
#pragma unroll
    for (int qq = 0; qq < 4; qq++) am[qq] = RowSum<double, kRowLanes64>(am[qq]), as[qq] = RowSum<double, kRowLanes64>(as[qq]);
    if (lane < kc) {
      const int ch = c0 + lane;
      q[(long)ch * N + bgi[r]] = w[2 * ch] * Pick(am, lane) + w[2 * ch + 1] * Pick(as, lane);
    }
  }
}

// q = A p for per-column operators in one launch: deep rows, then boundary rows.
__global__ __launch_bounds__(256) void K64ApplyW(int deepBlocks, int D, const int* __restrict__ deepLine, const int* __restrict__ lines,
                                                 const double* __restrict__ symM, const double* __restrict__ symS, int B,
                                                 const int* __restrict__ bgi, const int* __restrict__ mrpd, const int* __restrict__ mcold,
                                                 const double* __restrict__ mvald, const int* __restrict__ mrpb, const int* __restrict__ mcolb,
                                                 const double* __restrict__ mvalb, const int* __restrict__ srpd, const int* __restrict__ scold,
                                                 const double* __restrict__ svald, const int* __restrict__ srpb, const int* __restrict__ scolb,
                                                 const double* __restrict__ svalb, const double* __restrict__ p, int N, double* q, int K,
                                                 const double* __restrict__ w) {
  if ((int)blockIdx.x < deepBlocks) K64ApplyDeepWRows(XcdTile(blockIdx.x, deepBlocks), D, deepLine, lines, symM, symS, p, N, q, K, w);
  else
    K64ApplyBoundaryWRows(blockIdx.x - deepBlocks, B, bgi, mrpd, mcold, mvald, mrpb, mcolb, mvalb, srpd, scold, svald, srpb, scolb, svalb, p,
                          N, q, K, w);
}

// column dot products a_c . b_c: one partial sum per block (out[c][blockIdx.x]), no atomics, so
// that the flexible-CG scalars are reproducible run to run (blockIdx.y = column)
__global__ __launch_bounds__(256) void K64Dot(int N, const double* __restrict__ a, const double* __restrict__ b, double* out) {
  const int ch = blockIdx.y;
  double v = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) v += a[(long)ch * N + i] * b[(long)ch * N + i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __shared__ double part[4];
  if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[(long)ch * kDotBlocks + blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

__device__ __forceinline__ double DotSum(const double* s, int n) {
  double t = 0.0;
  for (int q = 0; q < n; q++) t += s[q];
  return t;
}

// beta_c = (z, q_prev) / (p_prev, q_prev); zero on the first step of a solve (flag set by the host)
__global__ void KFcgBeta(int K, int nb, const double* __restrict__ zq, const double* __restrict__ pqPrev, const int* __restrict__ first,
                         double* beta) {
  const int c = threadIdx.x;
  if (c >= K) return;
  beta[c] = (*first || !(pqPrev[c] > 0)) ? 0.0 : DotSum(zq + (long)c * kDotBlocks, nb) / pqPrev[c];
}

__global__ void KFcgDirection(long n, int N, const double* __restrict__ z, const double* __restrict__ beta, double* p) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double bt = beta[t / N];
  p[t] = bt == 0.0 ? z[t] : z[t] - bt * p[t];
}

// alpha_c = (r, p) / (p, q); keeps (p, q) for the next beta and clears the first-step flag
__global__ void KFcgAlpha(int K, int nb, const double* __restrict__ rp, const double* __restrict__ pq, double* alpha, double* pqPrev,
                          int* first) {
  const int c = threadIdx.x;
  if (c < K) {
    const double d = DotSum(pq + (long)c * kDotBlocks, nb);
    alpha[c] = d > 0 ? DotSum(rp + (long)c * kDotBlocks, nb) / d : 0.0;
    pqPrev[c] = d;
  }
  if (c == 0) *first = 0;
}

// x += alpha p, r -= alpha q; the new residual into the FP32 correction rhs (e zeroed), its norms
__global__ __launch_bounds__(256) void KFcgUpdate(int N, const int* __restrict__ g2d, const double* __restrict__ alpha,
                                                  const double* __restrict__ p, const double* __restrict__ q, double* x, double* r,
                                                  float* rhs0, float* e0, long Nd, double* norms) {
  const int ch = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0;
  if (i < N) {
    const long o = (long)ch * N + i;
    const double al = alpha[ch];
    x[o] += al * p[o];
    const double rv = r[o] - al * q[o];
    r[o] = rv;
    const long d = (long)ch * Nd + g2d[i];
    rhs0[d] = (float)rv;
    e0[d] = 0.0f;
    a = rv * rv;
  }
  BlockAddSquares(a, norms + (long)ch * kSlots);
}

// The fused sweeps swap a level's two correction buffers.  Level 0's correction is consumed (x += e,
// or z = e in flexible CG) at the end of a refinement step and zeroed for the next step; that zeroing
// goes to the buffer the level was allocated with, so that every step starts on the same buffer —
// a captured step replays with the pointers it was captured with, whatever the parity of its sweeps.
static void HomeLevel0(MixedSystem& ms) {
  DevMixLevel& d = ms.L[0];
  if (d.e2 && d.e != d.eHome) std::swap(d.e, d.e2);
}

// residual of the level-0 FP64 system into the FP32 correction rhs, squared norms into norms[0..K)
// bnorms (first residual of a solve): |b|^2 per column into bnorms[ch][slot] as well.
static void OuterResidual(DeviceSystem& ds, MixedSystem& ms, double* x, const double* b, int K, bool weighted, hipStream_t s,
                          double* r64 = nullptr, double* bnorms = nullptr, bool zeroGuess = false) {
  DevLevel& L0 = ds.L[0];
  DevMixLevel& d = ms.L[0];
  KMixZeroD<<<NBlk((long)K * kSlots, 256), 256, 0, s>>>(K * kSlots, ms.norms);
  if (zeroGuess) {  // x = 0: r = b, whatever the operator; refinement's first update writes x
    if (d.N > 0) K64ZeroGuess<<<NBlk(d.N, 256), 256, 0, s>>>(d.N, d.g2d, b, x, r64 != nullptr, d.rhs, d.e, d.Nd, ms.norms, bnorms, r64, K);
    return;
  }
  // deep rows: kDeepGroups x 256 per block; boundary rows: kRowLanes64 lanes each
  const int nt = 256, db = L0.nRaster > 0 ? (int)NBlk(L0.D, 256 * kDeepGroups) : 0;
  const int bb = L0.B > 0 ? (int)NBlk((long)L0.B * kRowLanes64, nt) : 0;
  if (db + bb == 0) return;
  if (weighted) {
    const DevLevel &M = ds.mass0, &S = ds.stiff0;
    K64ResidualW<<<db + bb, nt, 0, s>>>(db, L0.D, ms.deepLine, L0.raster, M.coefSym, S.coefSym, L0.B, L0.bgi, M.bd.rowptr, M.bd.col,
                                         M.bd.val, M.bb.rowptr, M.bb.col, M.bb.val, S.bd.rowptr, S.bd.col, S.bd.val, S.bb.rowptr, S.bb.col,
                                         S.bb.val, x, b, L0.N, d.g2d, d.rhs, d.e, d.Nd, ms.norms, K, ms.weights, r64, bnorms);
    return;
  }
  K64Residual<<<db + bb, nt, 0, s>>>(db, L0.D, ms.deepLine, L0.raster, L0.coefSym, L0.B, L0.bgi, L0.bd.rowptr, L0.bd.col, L0.bd.val,
                                      L0.bb.rowptr, L0.bb.col, L0.bb.val, x, b, L0.N, d.g2d, d.rhs, d.e, d.Nd, ms.norms, K, bnorms, ms.fresh);
}

// One flexible-CG step after the V-cycle left z = V(r) in the level-0 FP32 correction:
// beta = (z, q_prev)/(p_prev, q_prev), p = z - beta p, q = A p, alpha = (r, p)/(p, q),
// x += alpha p, r -= alpha q (all scalars stay on the device).
static void FcgStep(DeviceSystem& ds, MixedSystem& ms, double* x, int K, hipStream_t s) {
  DevLevel& L0 = ds.L[0];
  DevMixLevel& d = ms.L[0];
  const int N = d.N;
  const long nk = (long)N * K;
  const size_t slab = (size_t)ms.kcap * kDotBlocks;
  double *zq = ms.dots, *rp = ms.dots + slab, *pq = ms.dots + 2 * slab;
  const unsigned gb = std::min(NBlk(N, 256), (unsigned)kDotBlocks), cb = (unsigned)((K + 63) / 64 * 64);
  KMixGather<<<NBlk(nk, 256), 256, 0, s>>>(N, d.g2d, d.e, d.Nd, ms.z64, K);
  K64Dot<<<dim3(gb, K), 256, 0, s>>>(N, ms.z64, ms.q64, zq);
  KFcgBeta<<<1, cb, 0, s>>>(K, (int)gb, zq, ms.pqPrev, ms.first, ms.beta);
  KFcgDirection<<<NBlk(nk, 256), 256, 0, s>>>(nk, N, ms.z64, ms.beta, ms.p64);
  const DevLevel &M = ds.mass0, &S = ds.stiff0;
  const int db = L0.nRaster > 0 ? (int)NBlk(L0.D, 256) : 0;
  const int bb = L0.B > 0 ? (int)NBlk((long)L0.B * kRowLanes64, 256) : 0;
  if (db + bb > 0)
    K64ApplyW<<<db + bb, 256, 0, s>>>(db, L0.D, ms.deepLine, L0.raster, M.coefSym, S.coefSym, L0.B, L0.bgi, M.bd.rowptr, M.bd.col, M.bd.val,
                                      M.bb.rowptr, M.bb.col, M.bb.val, S.bd.rowptr, S.bd.col, S.bd.val, S.bb.rowptr, S.bb.col, S.bb.val,
                                      ms.p64, N, ms.q64, K, ms.weights);
  K64Dot<<<dim3(gb, K), 256, 0, s>>>(N, ms.r64, ms.p64, rp);
  K64Dot<<<dim3(gb, K), 256, 0, s>>>(N, ms.p64, ms.q64, pq);
  KFcgAlpha<<<1, cb, 0, s>>>(K, (int)gb, rp, pq, ms.alpha, ms.pqPrev, ms.first);
  KMixZeroD<<<NBlk((long)K * kSlots, 256), 256, 0, s>>>(K * kSlots, ms.norms);
  HomeLevel0(ms);  // e was gathered above; the update zeroes the buffer the next step starts on
  KFcgUpdate<<<dim3(NBlk(N, 256), K), 256, 0, s>>>(N, d.g2d, ms.alpha, ms.p64, ms.q64, x, ms.r64, d.rhs, d.e, d.Nd, ms.norms);
}

static void OuterUpdate(DeviceSystem& ds, MixedSystem& ms, double* x, int K, hipStream_t s) {
  DevMixLevel& d = ms.L[0];
  K64Update<<<NBlk(d.N, 256), 256, 0, s>>>(d.N, d.g2d, d.e, d.Nd, x, K, ms.fresh);
  (void)ds;
}

int SolveMixed(DeviceSystem& ds, MixedSystem& ms, double* x, const double* b, int K, double tol, int maxIter, double* relRes,
               hipStream_t s, bool useGraph, const double* weights, void (*between)(void*), void* ctx, bool zeroGuess) {
  if (K < 1 || K > ms.kcap) throw std::runtime_error("column count outside [1, capacity]");
  if (K > 1024) throw std::runtime_error("at most 1024 columns per solve");
  // weighted columns run flexible CG: their operators differ from the one the V-cycle was built
  // for, and plain refinement x += V(r) diverges once the weights are a few times apart
  const bool krylov = ms.krylov || weights != nullptr;
  std::vector<double> own;
  if (krylov && !weights) {  // flexible CG applies A through the level-0 mass and stiffness
    own.resize(2 * (size_t)K);
    for (int k = 0; k < K; k++) own[2 * k] = ms.screen, own[2 * k + 1] = ms.stiff;
    weights = own.data();
  }
  const bool weighted = weights != nullptr;
  ms.schwarzHalf = !krylov;  // FP16 Schwarz quads for plain refinement only (captured with the step)
  if (weighted) {
    for (int k = 0; k < 2 * K; k++)
      if (!(weights[k] >= 0) || !std::isfinite(weights[k])) throw std::runtime_error("operator weights must be finite and non-negative");
    for (int k = 0; k < K; k++)
      if (!(weights[2 * k] > 0)) throw std::runtime_error("the mass weight of every column must be positive");
    GDTP_HIP(hipMemcpyAsync(ms.weights, weights, 2 * (size_t)K * sizeof(double), hipMemcpyHostToDevice, s));
  }
  double* bNorms = ms.norms + (size_t)ms.kcap * kSlots;
  KMixZeroD<<<NBlk((long)K * kSlots, 256), 256, 0, s>>>(K * kSlots, bNorms);
  if (krylov) GDTP_HIP(hipMemsetD32Async((hipDeviceptr_t)ms.first, 1, 1, s));
  // refinement from a zero guess leaves x unwritten until its first update (x = 0 + e)
  const bool unwritten = zeroGuess && !krylov;
  GDTP_HIP(hipMemsetD32Async((hipDeviceptr_t)ms.fresh, unwritten ? 1 : 0, 1, s));
  HomeLevel0(ms);
  OuterResidual(ds, ms, x, b, K, weighted, s, krylov ? ms.r64 : nullptr, bNorms, zeroGuess);  // |b|^2 in the same pass
  GDTP_HIP(hipMemcpyAsync(ms.hostNorms, ms.norms, 2 * (size_t)ms.kcap * kSlots * sizeof(double), hipMemcpyDeviceToHost, s));
  GDTP_HIP(hipStreamSynchronize(s));
  auto slotSum = [&](const double* p) {
    double t = 0;
    for (int q = 0; q < kSlots; q++) t += p[q];
    return t;
  };
  std::vector<double> bn(K);
  for (int k = 0; k < K; k++) bn[k] = std::sqrt(std::max(slotSum(ms.hostNorms + ((size_t)ms.kcap + k) * kSlots), 1e-300));
  auto converged = [&]() {
    bool ok = true;
    for (int k = 0; k < K; k++) {
      relRes[k] = std::sqrt(slotSum(ms.hostNorms + (size_t)k * kSlots)) / bn[k];
      if (!(relRes[k] <= tol)) ok = false;
    }
    return ok;
  };
  if (converged() || maxIter <= 0) {
    if (unwritten) {  // no update will run: x = 0 as asked
      GDTP_HIP(hipMemsetAsync(x, 0, (size_t)K * ms.L[0].N * sizeof(double), s));
      GDTP_HIP(hipStreamSynchronize(s));
    }
    return 0;
  }
  // automatic: two V-cycles per residual for refinement on operators that FP32 represents well;
  // one for flexible CG (a product of two V-cycles preconditions it worse once the columns'
  // weights differ from the V-cycle's: 11 iterations of 2 cycles against 11 of 1 at a weight 3x off)
  int cycles = ms.cycles > 0 ? ms.cycles : (!krylov && ms.rowSumRatio >= kTwoCycleRatio ? 2 : 1);
  // one refinement step = V-cycles on the current residual, update, new residual
  auto step = [&](hipStream_t st, int cyc) {
    for (int c = 0; c < cyc; c++) MixedVCycle(ds, ms, K, st, c == 0);  // the residual step zeroed e
    if (krylov) {
      FcgStep(ds, ms, x, K, st);
    } else {
      OuterUpdate(ds, ms, x, K, st);
      HomeLevel0(ms);  // the residual zeroes the buffer the next step starts on
      OuterResidual(ds, ms, x, b, K, weighted, st);
    }
    GDTP_HIP(hipMemcpyAsync(ms.hostNorms, ms.norms, (size_t)K * kSlots * sizeof(double), hipMemcpyDeviceToHost, st));
  };
  // the captured step for (K, x, b, weighted, cycles): shape-static, replayed on the solver's stream
  auto graphFor = [&](int cyc) {
    MixedSystem::GraphEntry* ge = nullptr;
    for (auto& g : ms.graphs)
      if (g.K == K && g.x == x && g.b == b && g.weighted == weighted && g.cycles == cyc) ge = &g;
    if (!ge) {
      if (ms.graphs.size() >= kMaxGraphs) {
        auto old = std::min_element(ms.graphs.begin(), ms.graphs.end(), [](const auto& a, const auto& c) { return a.lastUse < c.lastUse; });
        if (old->exec) (void)hipGraphExecDestroy(old->exec);
        if (old->graph) (void)hipGraphDestroy(old->graph);
        ms.graphs.erase(old);
      }
      MixedSystem::GraphEntry g;
      GDTP_HIP(hipStreamBeginCapture(ms.stream, hipStreamCaptureModeThreadLocal));
      step(ms.stream, cyc);
      GDTP_HIP(hipStreamEndCapture(ms.stream, &g.graph));
      GDTP_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
      g.K = K, g.x = x, g.b = b, g.weighted = weighted, g.cycles = cyc;
      ms.graphs.push_back(g);
      ge = &ms.graphs.back();
    }
    ge->lastUse = ++ms.graphClock;
    return ge;
  };
  MixedSystem::GraphEntry* ge = useGraph ? graphFor(cycles) : nullptr;
  // (Measured and not kept: sizing the V-cycles of later steps to the per-cycle reduction the first
  // step showed, so that C2 would run 2 + 4 V-cycles on two FP64 residuals instead of 2 + 2 + 2 on
  // three.  Four V-cycles on one FP32 correction equation reduce its residual less than two and two
  // with a fresh FP64 residual between them, and C2 needed a third step: 7 V-cycles, 16.2 ms.)
  std::vector<double> prev(relRes, relRes + K);
  int it = 0;
  while (it < maxIter) {
    if (useGraph) {
      GDTP_HIP(hipGraphLaunch(ge->exec, ms.stream));
      GDTP_HIP(hipStreamSynchronize(ms.stream));
    } else {
      step(s, cycles);
      GDTP_HIP(hipStreamSynchronize(s));
    }
    it++;
    if (converged()) break;
    // Safety net for several V-cycles per iteration: a residual that grows means the FP32
    // operator misrepresents a near-null mode; continue with one V-cycle per FP64 residual.
    bool grew = false;
    for (int k = 0; k < K; k++) grew = grew || !(relRes[k] <= prev[k]);
    if (grew && cycles > 1) {
      cycles = 1;
      if (useGraph) ge = graphFor(cycles);
    }
    prev.assign(relRes, relRes + K);
    if (between && it < maxIter) between(ctx);
  }
  return it;
}

}  // namespace gdtp
