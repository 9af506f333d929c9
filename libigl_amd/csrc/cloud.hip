// cloud.hip -- point-cloud fast winding number on gfx950.
//
// Octree (the reference's igl::octree, octree.cpp:7-169, cell for cell):
//   k_oct_keys    each point's octant path, 42 levels, computed with the
//                 reference's float centre recursion (child centre = parent
//                 centre +- width/4, octant bits x>=c | 2(y>=c) | 4(z>=c)),
//                 packed as two 63-bit words (3 bits per level, level 0 first)
//   radix sort    stable LDS radix sort on (low word, then high word): points
//                 of one cell become contiguous and keep ascending indices,
//                 which is the order of the reference's point_indices lists
//   k_oct_lcp     common-prefix length (in levels) of neighbouring keys, and
//                 the number of split cells each neighbour boundary opens
//   scan          -> the depth-first rank of every split cell: the reference
//                 appends the 8 children of the k-th split cell (in depth-first
//                 order) as cells 1+8k .. 8+8k, so rank gives the cell ids
//   k_oct_level   one launch per level: split each cell's point range by the
//                 next octant digit, write CH, centres, widths, point ranges,
//                 and queue the children that hold two or more points
// Cells deeper than 42 levels are not split (the reference keeps splitting up
// to depth 30000, which only matters for points equal to ~1 float ulp).
//
// Precompute (fast_winding_number.cpp:8-124): per non-empty cell, CM = area-
// weighted centre, R = max |p - CM|, EC = [sum A n, sum A (p-CM) n^T,
// sum 1/2 (p-CM)_k A (p-CM) n^T], accumulated in FP64 (a thread per small
// cell, a wavefront per large one) and stored in FP32.
//
// Evaluation (fast_winding_number.cpp:126-311): the reference's near lists
// become a rope over the non-empty non-root cells in depth-first order: a
// non-leaf cell with |CM-q| > beta R adds its Taylor expansion and jumps over
// its subtree, otherwise the walk steps into its children; a leaf adds the
// direct dipole sum of its points (which the reference does whether the leaf
// is near or far).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_math.h"
#include "gwn_internal.h"

namespace gwn {

constexpr int OCT_LEVELS = 42;  // octant digits per point (two words of 21)
constexpr int OCT_MAX_SPLIT_DEPTH = OCT_LEVELS - 1;
constexpr uint32_t REC_LEAF = 0x80000000u;
constexpr double CL_PI = 3.1415926535897932384626433832795;

__global__ void k_oct_keys(const float* __restrict__ P, int64_t n, float cx, float cy, float cz, float w0,
                           uint64_t* __restrict__ hi, uint64_t* __restrict__ lo, uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
  float c0 = cx, c1 = cy, c2 = cz, w = w0;
  uint64_t kh = 0, kl = 0;
  for (int d = 0; d < OCT_LEVELS; ++d) {
    uint32_t o = 0;
    if (x >= c0) o += 1;
    if (y >= c1) o += 2;
    if (z >= c2) o += 4;
    if (d < 21)
      kh |= (uint64_t)o << (60 - 3 * d);
    else
      kl |= (uint64_t)o << (60 - 3 * (d - 21));
    const float h = w / 2;   // child width
    const float hh = h / 2;  // child centre offset
    c0 = c0 + ((o & 1) ? hh : -hh);
    c1 = c1 + ((o & 2) ? hh : -hh);
    c2 = c2 + ((o & 4) ? hh : -hh);
    w = h;
  }
  hi[i] = kh;
  lo[i] = kl;
  idx[i] = (uint32_t)i;
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint32_t* __restrict__ perm, int64_t n,
                             uint64_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__device__ __forceinline__ int oct_lcp(uint64_t h0, uint64_t l0, uint64_t h1, uint64_t l1) {
  const uint64_t x = h0 ^ h1;
  if (x) return (__clzll(x) - 1) / 3;
  const uint64_t y = l0 ^ l1;
  if (y) return 21 + (__clzll(y) - 1) / 3;
  return OCT_LEVELS;
}

// lcp[a] for boundary a (between sorted points a and a+1), a in [0, n-1);
// opens[a] = number of split cells whose first point is a.
__global__ void k_oct_lcp(const uint64_t* __restrict__ hi, const uint64_t* __restrict__ lo, int64_t n,
                          int32_t* __restrict__ lcp, uint32_t* __restrict__ opens) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n - 1) return;
  const int cur = oct_lcp(hi[a], lo[a], hi[a + 1], lo[a + 1]);
  const int prev = a == 0 ? -1 : oct_lcp(hi[a - 1], lo[a - 1], hi[a], lo[a]);
  lcp[a] = cur;
  const int top = cur < OCT_MAX_SPLIT_DEPTH ? cur : OCT_MAX_SPLIT_DEPTH;
  opens[a] = top > prev ? (uint32_t)(top - prev) : 0u;
}

__device__ __forceinline__ uint32_t oct_digit(const uint64_t* hi, const uint64_t* lo, int64_t k, int d) {
  return d < 21 ? (uint32_t)((hi[k] >> (60 - 3 * d)) & 7u) : (uint32_t)((lo[k] >> (60 - 3 * (d - 21))) & 7u);
}

struct SplitCell {
  int32_t cell;
  int32_t depth;
  int64_t start, end;
  uint32_t rank;
  uint32_t pad;
};

__device__ __forceinline__ uint32_t split_rank(const int32_t* lcp, const uint32_t* open_scan, int64_t s, int d) {
  const int prev = s == 0 ? -1 : lcp[s - 1];
  return open_scan[s] + (uint32_t)(d - prev - 1);
}

__global__ void k_oct_level(const SplitCell* __restrict__ cur, uint32_t ncur, const uint64_t* __restrict__ hi,
                            const uint64_t* __restrict__ lo, const int32_t* __restrict__ lcp,
                            const uint32_t* __restrict__ open_scan, int32_t* __restrict__ CH, float* __restrict__ CN,
                            float* __restrict__ Wd, int64_t* __restrict__ cstart, int32_t* __restrict__ ccount,
                            SplitCell* __restrict__ next, uint32_t* __restrict__ nnext, uint32_t next_cap,
                            int64_t ncells, uint32_t* __restrict__ err) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncur) return;
  const SplitCell sc = cur[t];
  const float c0 = CN[3 * (int64_t)sc.cell], c1 = CN[3 * (int64_t)sc.cell + 1], c2 = CN[3 * (int64_t)sc.cell + 2];
  const float h = Wd[sc.cell] / 2;
  const float hh = h / 2;
  int64_t s = sc.start;
  for (uint32_t o = 0; o < 8; ++o) {
    // first point of the range whose digit at this depth exceeds o
    int64_t a = s, b = sc.end;
    while (a < b) {
      const int64_t m = (a + b) >> 1;
      if (oct_digit(hi, lo, m, sc.depth) <= o)
        a = m + 1;
      else
        b = m;
    }
    const int64_t e = a;
    const int64_t child64 = 1 + 8 * (int64_t)sc.rank + o;
    if (child64 >= ncells) {  // inconsistent ranks: report, never write out of range
      atomicOr(err, 1u);
      return;
    }
    const int32_t child = (int32_t)child64;
    CH[8 * (int64_t)sc.cell + o] = child;
    CN[3 * (int64_t)child] = c0 + ((o & 1) ? hh : -hh);
    CN[3 * (int64_t)child + 1] = c1 + ((o & 2) ? hh : -hh);
    CN[3 * (int64_t)child + 2] = c2 + ((o & 4) ? hh : -hh);
    Wd[child] = h;
    cstart[child] = s;
    ccount[child] = (int32_t)(e - s);
    if (e - s >= 2 && sc.depth + 1 <= OCT_MAX_SPLIT_DEPTH) {
      SplitCell nc;
      nc.cell = child;
      nc.depth = sc.depth + 1;
      nc.start = s;
      nc.end = e;
      nc.rank = split_rank(lcp, open_scan, s, sc.depth + 1);
      nc.pad = 0;
      const uint32_t slot = atomicAdd(nnext, 1u);
      if (slot < next_cap)
        next[slot] = nc;
      else
        atomicOr(err, 2u);
    }
    s = e;
  }
}

__global__ void k_split_counts(const SplitCell* __restrict__ lvl, uint32_t m, uint32_t* __restrict__ cnt) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) cnt[j] = (uint32_t)(lvl[j].end - lvl[j].start);
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_permute_points(const float* __restrict__ P, const float* __restrict__ N, const float* __restrict__ A,
                                 const uint32_t* __restrict__ perm, int64_t n, float* __restrict__ Ps,
                                 float* __restrict__ NAs, int32_t* __restrict__ perm_out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = perm[k];
  const float a = A[i];
  for (int d = 0; d < 3; ++d) {
    Ps[3 * k + d] = P[3 * i + d];
    NAs[3 * k + d] = N[3 * i + d] * a;  // N.row * A as the reference forms it
  }
  perm_out[k] = (int32_t)i;
}

// ---------------- precompute ----------------
// Centre of mass of split cells, summed as the reference sums it: FP32,
// sequentially, over the cell's points in ascending index order
// (fast_winding_number.cpp:60-73).  Per level, the (cell, point) pairs are
// sorted by (cell ordinal, point index), the terms are laid out in that
// order, and each cell's four sums are folded in it: a wavefront per cell on
// levels of few (large) cells, a thread per cell on levels of many.
// Block (j, y): cell j's pairs y * 64 + lane, stepping by gridDim.y * 64, so
// that a level of few large cells still spreads over many blocks.
__global__ void k_lvl_emit(const SplitCell* __restrict__ lvl, const uint32_t* __restrict__ off,
                           const int32_t* __restrict__ perm, uint64_t* __restrict__ keys) {
  const uint32_t j = blockIdx.x;
  const SplitCell sc = lvl[j];
  const int64_t cnt = sc.end - sc.start;
  const uint32_t o = off[j];
  const int64_t step = (int64_t)gridDim.y * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; t < cnt; t += step)
    keys[o + t] = ((uint64_t)j << 32) | (uint32_t)perm[sc.start + t];
}

// The terms the reference adds for a cell, in key order: a, a*x, a*y, a*z
// (FP32 products, fast_winding_number.cpp:64-66).
__global__ void k_lvl_terms(int64_t T, const uint64_t* __restrict__ keys, const float* __restrict__ P,
                            const float* __restrict__ A, float4* __restrict__ terms) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int64_t i = (int64_t)(keys[t] & 0xFFFFFFFFull);
  const float a = A[i];
  terms[t] = make_float4(a, a * P[3 * i], a * P[3 * i + 1], a * P[3 * i + 2]);
}

// One thread per cell (levels of many small cells).
__global__ void k_lvl_fold(const SplitCell* __restrict__ lvl, uint32_t ncell, const uint32_t* __restrict__ off,
                           const float4* __restrict__ terms, float* __restrict__ CM) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncell) return;
  const SplitCell sc = lvl[j];
  const int64_t cnt = sc.end - sc.start;
  const float4* v = terms + off[j];
  float at = 0.0f, m0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
  for (int64_t t = 0; t < cnt; ++t) {
    const float4 q = v[t];
    at += q.x;
    m0 += q.y;
    m1 += q.z;
    m2 += q.w;
  }
  CM[3 * (int64_t)sc.cell] = m0 / at;
  CM[3 * (int64_t)sc.cell + 1] = m1 / at;
  CM[3 * (int64_t)sc.cell + 2] = m2 / at;
}

constexpr int FOLD_CHUNK = 512;   // terms staged in LDS per round
constexpr int FOLD_GROUP = 16;    // float4 LDS reads in flight per lane

// s + b[0] + b[1] + ... + b[n-1], added in that order (FP32).  The lane reads
// 64 of its terms from LDS before adding them, so the LDS latency is paid
// once per 64 dependent adds, not once per 4.
__device__ __forceinline__ float fold_in_order(const float* b, int n, float s) {
  const float4* b4 = reinterpret_cast<const float4*>(b);
  const int n4 = n >> 2;
  int g = 0;
  for (; g + FOLD_GROUP <= n4; g += FOLD_GROUP) {
    float4 q[FOLD_GROUP];
#pragma unroll
    for (int k = 0; k < FOLD_GROUP; ++k) q[k] = b4[g + k];
#pragma unroll
    for (int k = 0; k < FOLD_GROUP; ++k) {
      s += q[k].x;
      s += q[k].y;
      s += q[k].z;
      s += q[k].w;
    }
  }
  for (int e = 4 * g; e < n; ++e) s += b[e];
  return s;
}

// One wavefront per cell (levels of few, large cells).  The lanes stage
// FOLD_CHUNK terms at a time in LDS, the next FOLD_CHUNK already in flight,
// and lanes 0-3 each fold one of the four running sums in order: the same
// sequential FP32 sums, one add per term and lane.
__global__ __launch_bounds__(64) void k_lvl_fold_wave(const SplitCell* __restrict__ lvl,
                                                      const uint32_t* __restrict__ off,
                                                      const float4* __restrict__ terms, float* __restrict__ CM) {
  __shared__ __align__(16) float buf[4][FOLD_CHUNK];
  const SplitCell sc = lvl[blockIdx.x];
  const int64_t cnt = sc.end - sc.start;
  const float4* v = terms + off[blockIdx.x];
  const int lane = threadIdx.x;
  constexpr int PER = FOLD_CHUNK / 64;
  float4 r[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u)
    if (lane + 64 * u < cnt) r[u] = v[lane + 64 * u];
  float s = 0.0f;
  for (int64_t base = 0; base < cnt; base += FOLD_CHUNK) {
    const int n = (int)(cnt - base < FOLD_CHUNK ? cnt - base : FOLD_CHUNK);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = lane + 64 * u;
      if (e < n) {
        buf[0][e] = r[u].x;
        buf[1][e] = r[u].y;
        buf[2][e] = r[u].z;
        buf[3][e] = r[u].w;
      }
    }
    __syncthreads();
    const int64_t nb = base + FOLD_CHUNK;
#pragma unroll
    for (int u = 0; u < PER; ++u)
      if (nb + lane + 64 * u < cnt) r[u] = v[nb + lane + 64 * u];
    if (lane < 4) s = fold_in_order(buf[lane], n, s);
    __syncthreads();
  }
  const float at = __shfl(s, 0);
  if (lane >= 1 && lane <= 3) CM[3 * (int64_t)sc.cell + lane - 1] = s / at;
}

struct CellSums {
  double area, ap[3], an[3];
};

template <int NT>
__device__ __forceinline__ void cell_moments_serial(const float* __restrict__ P, const float* __restrict__ N,
                                                    const float* __restrict__ A, const int32_t* __restrict__ perm,
                                                    int64_t s, int32_t cnt, bool split, float* __restrict__ cm,
                                                    float* __restrict__ r, float* __restrict__ ec) {
  double an[3] = {0, 0, 0};
  float at = 0.0f, mc[3] = {0.0f, 0.0f, 0.0f};
  for (int32_t j = 0; j < cnt; ++j) {
    const int64_t p = perm[s + j];
    const float af = A[p];
    at += af;
    for (int d = 0; d < 3; ++d) {
      mc[d] += af * P[3 * p + d];
      an[d] += (double)af * (double)N[3 * p + d];
    }
  }
  float c[3];
  for (int d = 0; d < 3; ++d) c[d] = split ? cm[d] : mc[d] / at;
  double e[NT > 3 ? NT - 3 : 1];
  for (int k = 0; k < NT - 3; ++k) e[k] = 0;
  float rmax = 0.0f;
  for (int32_t j = 0; j < cnt; ++j) {
    const int64_t p = perm[s + j];
    const double a = A[p];
    float q[3];
    for (int d = 0; d < 3; ++d) q[d] = P[3 * p + d] - c[d];
    const float nrm = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    rmax = nrm > rmax ? nrm : rmax;
    if (NT >= 12) {
      const double n3[3] = {N[3 * p], N[3 * p + 1], N[3 * p + 2]};
      for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) {
          const double t = a * (double)q[u] * n3[v];
          e[u + 3 * v] += t;
          if (NT == 39)
            for (int k = 0; k < 3; ++k) e[9 + 9 * k + u + 3 * v] += 0.5 * (double)q[k] * t;
        }
    }
  }
  for (int d = 0; d < 3; ++d) {
    cm[d] = c[d];
    ec[d] = (float)an[d];
  }
  for (int k = 0; k < NT - 3; ++k) ec[3 + k] = (float)e[k];
  *r = rmax;
}

// one thread per small cell
template <int NT>
__global__ void k_cell_small(const int32_t* __restrict__ cells, uint32_t count, const int32_t* __restrict__ CH,
                             bool fold_cm, const int64_t* __restrict__ cstart,
                             const int32_t* __restrict__ ccount, const float* __restrict__ P,
                             const float* __restrict__ N, const float* __restrict__ A,
                             const int32_t* __restrict__ perm, float* __restrict__ CM, float* __restrict__ R,
                             float* __restrict__ EC) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t c = cells[t];
  cell_moments_serial<NT>(P, N, A, perm, cstart[c], ccount[c], fold_cm && CH[8 * (int64_t)c] >= 0, CM + 3 * (int64_t)c, R + c,
                          EC + (int64_t)NT * c);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Centre of a large cell that is not a split cell of this library's octree
// (an unsplittable leaf, or any cell of a caller's octree): the sequential
// FP32 sums over its list in list order, folded as k_lvl_fold_wave folds them
// (one wavefront per cell; split cells already have theirs).
__global__ __launch_bounds__(64) void k_cm_fold_wave(const int32_t* __restrict__ cells,
                                                     const int32_t* __restrict__ CH, bool fold_cm,
                                                     const int64_t* __restrict__ cstart,
                                                     const int32_t* __restrict__ ccount, const float* __restrict__ P,
                                                     const float* __restrict__ A, const int32_t* __restrict__ perm,
                                                     float* __restrict__ CM) {
  __shared__ __align__(16) float buf[4][FOLD_CHUNK];
  const int32_t c = cells[blockIdx.x];
  if (fold_cm && CH[8 * (int64_t)c] >= 0) return;  // uniform over the block
  const int64_t s0 = cstart[c];
  const int64_t cnt = ccount[c];
  const int lane = threadIdx.x;
  constexpr int PER = FOLD_CHUNK / 64;
  float4 r[PER];
  auto gather = [&](int64_t base) {
#pragma unroll
    for (int u = 0; u < PER; ++u)
      if (base + lane + 64 * u < cnt) {
        const int64_t p = perm[s0 + base + lane + 64 * u];
        const float a = A[p];
        r[u] = make_float4(a, a * P[3 * p], a * P[3 * p + 1], a * P[3 * p + 2]);
      }
  };
  gather(0);
  float s = 0.0f;
  for (int64_t base = 0; base < cnt; base += FOLD_CHUNK) {
    const int n = (int)(cnt - base < FOLD_CHUNK ? cnt - base : FOLD_CHUNK);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = lane + 64 * u;
      if (e < n) {
        buf[0][e] = r[u].x;
        buf[1][e] = r[u].y;
        buf[2][e] = r[u].z;
        buf[3][e] = r[u].w;
      }
    }
    __syncthreads();
    gather(base + FOLD_CHUNK);
    if (lane < 4) s = fold_in_order(buf[lane], n, s);
    __syncthreads();
  }
  const float at = __shfl(s, 0);
  if (lane >= 1 && lane <= 3) CM[3 * (int64_t)c + lane - 1] = s / at;
}

// Moments of a large cell, one wavefront per segment of at most
// MOMENT_SEGMENT of its points (the root alone spreads over many CUs): FP64
// partial sums of A n, A q n^T, 1/2 q_k A q n^T and the largest |q| (q = p -
// CM), written to part[item * NT .. ].
constexpr int MOMENT_SEGMENT = 4096;
template <int NT>
__global__ __launch_bounds__(256) void k_cell_large(const int32_t* __restrict__ item_cell,
                                                    const int32_t* __restrict__ item_j0, uint32_t nitems,
                                                    const int64_t* __restrict__ cstart,
                                                    const int32_t* __restrict__ ccount, const float* __restrict__ P,
                                                    const float* __restrict__ N, const float* __restrict__ A,
                                                    const int32_t* __restrict__ perm, const float* __restrict__ CM,
                                                    double* __restrict__ part) {
  const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (w >= nitems) return;
  const int32_t c = item_cell[w];
  const int64_t s = cstart[c];
  const int32_t j0 = item_j0[w];
  const int32_t j1 = j0 + MOMENT_SEGMENT < ccount[c] ? j0 + MOMENT_SEGMENT : ccount[c];
  const float c3[3] = {CM[3 * (int64_t)c], CM[3 * (int64_t)c + 1], CM[3 * (int64_t)c + 2]};
  double an0 = 0, an1 = 0, an2 = 0;
  double e[NT > 3 ? NT - 3 : 1];
  for (int k = 0; k < NT - 3; ++k) e[k] = 0;
  float rmax = 0.0f;
  for (int32_t j = j0 + lane; j < j1; j += 64) {
    const int64_t p = perm[s + j];
    const double a = A[p];
    const double n3[3] = {N[3 * p], N[3 * p + 1], N[3 * p + 2]};
    an0 += a * n3[0];
    an1 += a * n3[1];
    an2 += a * n3[2];
    float q[3];
    for (int d = 0; d < 3; ++d) q[d] = P[3 * p + d] - c3[d];
    const float nrm = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    rmax = nrm > rmax ? nrm : rmax;
    if (NT >= 12) {
      for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) {
          const double t = a * (double)q[u] * n3[v];
          e[u + 3 * v] += t;
          if (NT == 39)
            for (int k = 0; k < 3; ++k) e[9 + 9 * k + u + 3 * v] += 0.5 * (double)q[k] * t;
        }
    }
  }
  an0 = wave_sum(an0);
  an1 = wave_sum(an1);
  an2 = wave_sum(an2);
  rmax = wave_max(rmax);
  for (int k = 0; k < NT - 3; ++k) e[k] = wave_sum(e[k]);
  if (lane == 0) {
    double* o = part + (int64_t)(NT + 1) * w;
    o[0] = an0;
    o[1] = an1;
    o[2] = an2;
    for (int k = 0; k < NT - 3; ++k) o[3 + k] = e[k];
    o[NT] = rmax;
  }
}

// A large cell's segments summed in segment order; EC and R in FP32.
template <int NT>
__global__ void k_cell_combine(const int32_t* __restrict__ cells, uint32_t count, const uint32_t* __restrict__ item_off,
                               const double* __restrict__ part, float* __restrict__ R, float* __restrict__ EC) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t c = cells[t];
  double sum[NT];
  for (int k = 0; k < NT; ++k) sum[k] = 0.0;
  double rmax = 0.0;
  for (uint32_t it = item_off[t]; it < item_off[t + 1]; ++it) {
    const double* o = part + (int64_t)(NT + 1) * it;
    for (int k = 0; k < NT; ++k) sum[k] += o[k];
    rmax = o[NT] > rmax ? o[NT] : rmax;
  }
  R[c] = (float)rmax;
  for (int k = 0; k < NT; ++k) EC[(int64_t)NT * c + k] = (float)sum[k];
}

__global__ void k_large_counts(const int32_t* __restrict__ cells, uint32_t count, const int32_t* __restrict__ ccount,
                               int32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) out[t] = ccount[cells[t]];
}

__global__ void k_cell_empty(int64_t ncells, const int32_t* __restrict__ ccount, int nt, float* __restrict__ CM,
                             float* __restrict__ R, float* __restrict__ EC) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells || ccount[c] > 0) return;
  for (int d = 0; d < 3; ++d) CM[3 * c + d] = __int_as_float(0x7FC00000);  // NaN, as the reference
  R[c] = 0.0f;
  for (int k = 0; k < nt; ++k) EC[nt * c + k] = 0.0f;
}

// classify non-empty cells into small / large lists: flags, scans, scatter
__global__ void k_cell_flags(int64_t ncells, const int32_t* __restrict__ ccount, uint32_t* __restrict__ fs,
                             uint32_t* __restrict__ fl) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int32_t k = ccount[c];
  fs[c] = (k > 0 && k <= 32) ? 1u : 0u;
  fl[c] = k > 32 ? 1u : 0u;
}

__global__ void k_cell_lists(int64_t ncells, const uint32_t* __restrict__ fs, const uint32_t* __restrict__ ps,
                             const uint32_t* __restrict__ fl, const uint32_t* __restrict__ pl,
                             int32_t* __restrict__ small, int32_t* __restrict__ large) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  if (fs[c]) small[ps[c]] = (int32_t)c;
  if (fl[c]) large[pl[c]] = (int32_t)c;
}

// ---------------- rope over non-empty non-root cells ----------------
// bottom-up: rc[cell] = records in the subtree of a non-empty cell (itself included)
__global__ void k_rec_count(const SplitCell* __restrict__ lvl, uint32_t n, const int32_t* __restrict__ CH,
                            const int32_t* __restrict__ ccount, uint32_t* __restrict__ rc) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t c = lvl[t].cell;
  uint32_t s = 1;
  for (int o = 0; o < 8; ++o) {
    const int32_t ch = CH[8 * (int64_t)c + o];
    if (ccount[ch] <= 0) continue;
    s += (CH[8 * (int64_t)ch] >= 0) ? rc[ch] : 1u;
  }
  rc[c] = s;
}

// top-down: record positions of the children of each split cell
__global__ void k_rec_pos(const SplitCell* __restrict__ lvl, uint32_t n, const int32_t* __restrict__ CH,
                          const int32_t* __restrict__ ccount, const uint32_t* __restrict__ rc,
                          uint32_t* __restrict__ pos) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t c = lvl[t].cell;
  if (c != 0 && pos[c] == 0xFFFFFFFFu) return;  // not reached (an empty ancestor)
  uint32_t p = (c == 0) ? 0u : pos[c] + 1;
  for (int o = 0; o < 8; ++o) {
    const int32_t ch = CH[8 * (int64_t)c + o];
    if (ccount[ch] <= 0) continue;
    pos[ch] = p;
    p += (CH[8 * (int64_t)ch] >= 0) ? rc[ch] : 1u;
  }
}

__global__ void k_rec_write(int64_t ncells, const int32_t* __restrict__ CH, const int32_t* __restrict__ ccount,
                            const int64_t* __restrict__ cstart, const uint32_t* __restrict__ rc,
                            const uint32_t* __restrict__ pos, const float* __restrict__ CM,
                            const float* __restrict__ R, uint32_t nrec, float4* __restrict__ head,
                            uint2* __restrict__ meta) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0 || c >= ncells || ccount[c] <= 0) return;
  const uint32_t p = pos[c];
  if (p >= nrec) return;
  head[p] = make_float4(CM[3 * c], CM[3 * c + 1], CM[3 * c + 2], R[c]);
  if (CH[8 * c] >= 0)
    meta[p] = make_uint2(p + rc[c], (uint32_t)c);
  else
    meta[p] = make_uint2(REC_LEAF | (uint32_t)cstart[c], (uint32_t)ccount[c]);
}

// The expansion of each internal record as one polynomial in the direction
// to the query (far_field_poly, device_math.h), formed in FP64 from the cell's
// FP32 EC and rounded once.  With l = CM - q, x = l/|l| and m = 1/|l|, the
// terms of the reference's expansion (fast_winding_number.cpp:174-228) are
//   order 0: m^2 (x.E0) / 4pi
//   order 1: m^3 sum_uv (delta_uv - 3 x_u x_v) E1[u][v] / 4pi
//          = m^3 x^T (tr(E1) I - 3 E1) x / 4pi                  (|x| = 1)
//   order 2: m^4 [15 sum_iuv x_i x_u x_v E2[i][u][v]
//                 - 3 sum_iuv (delta_uv x_i + delta_ui x_v + delta_vi x_u) E2[i][u][v]] / 4pi
//          = m^4 [15 T(x) - 3 (x.s)|x|^2] / 4pi,
//     s_k = sum_u E2[k][u][u] + sum_i E2[i][i][k] + sum_i E2[i][k][i],
// with E1[u][v] = EC[3 + u + 3v] and E2[i][u][v] = EC[12 + 9i + u + 3v].  This
// is exact algebra; only the rounding of the far field changes (~1e-7
// relative), and the near/far decisions do not use it.
template <int NT>
__global__ void k_rec_poly(uint32_t nrec, const uint2* __restrict__ meta, const float* __restrict__ EC,
                           float4* __restrict__ poly) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nrec) return;
  const uint2 m = meta[p];
  if (m.x & REC_LEAF) return;  // leaf records sum their points; no expansion
  const float* ec = EC + (int64_t)NT * m.y;
  const double k = 1.0 / (4.0 * CL_PI);
  double b[3], q[6] = {0, 0, 0, 0, 0, 0}, g[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int u = 0; u < 3; ++u) b[u] = (double)ec[u] * k;
  if (NT > 3) {
    double E1[3][3];
    for (int u = 0; u < 3; ++u)
      for (int v = 0; v < 3; ++v) E1[u][v] = ec[3 + u + 3 * v];
    const double tr = E1[0][0] + E1[1][1] + E1[2][2];
    // c.x c.y c.z c.w d.x d.y = xx yy zz xy yz xz
    q[0] = (tr - 3.0 * E1[0][0]) * k;
    q[1] = (tr - 3.0 * E1[1][1]) * k;
    q[2] = (tr - 3.0 * E1[2][2]) * k;
    q[3] = -3.0 * (E1[0][1] + E1[1][0]) * k;
    q[4] = -3.0 * (E1[1][2] + E1[2][1]) * k;
    q[5] = -3.0 * (E1[0][2] + E1[2][0]) * k;
  }
  if (NT > 12) {
    // cubic monomials by exponents (ex, ey, ez), in the order of g[]:
    // x^3 y^3 z^3 x^2y x^2z xy^2 y^2z xz^2 yz^2 xyz
    auto slot = [](int ex, int ey, int ez) {
      if (ex == 3) return 0;
      if (ey == 3) return 1;
      if (ez == 3) return 2;
      if (ex == 2) return ey ? 3 : 4;
      if (ey == 2) return ex ? 5 : 6;
      if (ez == 2) return ex ? 7 : 8;
      return 9;
    };
    double sv[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
      for (int u = 0; u < 3; ++u)
        for (int v = 0; v < 3; ++v) {
          const double e = ec[12 + 9 * i + u + 3 * v];
          int ex[3] = {0, 0, 0};
          ++ex[i];
          ++ex[u];
          ++ex[v];
          g[slot(ex[0], ex[1], ex[2])] += 15.0 * e;
          if (u == v) sv[i] += e;
          if (u == i) sv[v] += e;
          if (v == i) sv[u] += e;
        }
    // -3 (x.s)(x^2 + y^2 + z^2)
    g[0] -= 3.0 * sv[0];  // x^3
    g[5] -= 3.0 * sv[0];  // x y^2
    g[7] -= 3.0 * sv[0];  // x z^2
    g[3] -= 3.0 * sv[1];  // x^2 y
    g[1] -= 3.0 * sv[1];  // y^3
    g[8] -= 3.0 * sv[1];  // y z^2
    g[4] -= 3.0 * sv[2];  // x^2 z
    g[6] -= 3.0 * sv[2];  // y^2 z
    g[2] -= 3.0 * sv[2];  // z^3
    for (int j = 0; j < 10; ++j) g[j] *= k;
  }
  // record layout of far_field_poly: b, c, d, e, f
  float4* r = poly + 5 * (size_t)p;
  r[0] = make_float4((float)b[0], (float)b[1], (float)b[2], 0.0f);
  r[1] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
  // d = (yz, xz, x^3, y^3), e = (z^3, x^2y, x^2z, xy^2), f = (y^2z, xz^2, yz^2, xyz)
  r[2] = make_float4((float)q[4], (float)q[5], (float)g[0], (float)g[1]);
  r[3] = make_float4((float)g[2], (float)g[3], (float)g[4], (float)g[5]);
  r[4] = make_float4((float)g[6], (float)g[7], (float)g[8], (float)g[9]);
}

// ---------------- evaluation ----------------
__device__ __forceinline__ float dipole(float lx, float ly, float lz, float ax, float ay, float az) {
  const float PI_4 = (float)(4.0 * CL_PI);
  const float n = sqrtf(lx * lx + ly * ly + lz * lz);
  if (n == 0.0f) return 0.5f;
  return (lx * ax + ly * ay + lz * az) / (PI_4 * (n * n * n));
}

// Wave-coherent rope walk (the scheme of the soup traversal): each lane keeps
// `skip`, the next record it needs; the wave walks one shared index and a lane
// takes part in a record when the index reaches its skip.  The wave steps to
// i+1 when some lane enters the cell (or at a leaf) and to the cell's skip
// otherwise; every other lane needs a record at or beyond it, the subtrees of
// the rope being nested intervals.  Each lane so visits the records of its own
// walk in order (results independent of the other queries in its wave) while
// every record, leaf point and coefficient a wave reads is the same for all
// its lanes and is fetched once through the scalar cache.
template <int NT>
__global__ __launch_bounds__(256) void k_cloud_traverse(const float4* __restrict__ head, const uint2* __restrict__ meta,
                                                        uint32_t nrec, const float4* __restrict__ poly,
                                                        const float* __restrict__ Ps, const float* __restrict__ NAs,
                                                        int64_t npts, bool root_leaf, const float* __restrict__ Q,
                                                        const uint32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ order, int64_t nq, double beta,
                                                        float* __restrict__ WN) {
  // order (when given): the logical block this block takes (costliest first)
  const int64_t t = (int64_t)(order ? order[xcd_block()] : xcd_block()) * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const int64_t qi = perm ? (int64_t)perm[t] : t;
  const float qx = Q[3 * qi], qy = Q[3 * qi + 1], qz = Q[3 * qi + 2];
  float wn = 0.0f;
  if (root_leaf) {
    for (int64_t j = 0; j < npts; ++j)
      wn += dipole(Ps[3 * j] - qx, Ps[3 * j + 1] - qy, Ps[3 * j + 2] - qz, NAs[3 * j], NAs[3 * j + 1], NAs[3 * j + 2]);
  } else {
    uint32_t i = 0, skip = 0;
    while (i < nrec) {
      const uint2 m = meta[i];
      const float4 h = head[i];
      const float4* r = poly + 5 * (size_t)i;  // in bounds for every record; unused on a leaf
      const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4];
      // the record's meta, head and polynomial in one scalar-cache round trip
      // (the compiler would otherwise wait for meta, then for the decision,
      // before requesting the rest; measured 1.84 -> 1.72 ms on 4.1M queries)
      asm volatile("" ::"s"(m.x), "s"(m.y), "s"(h.x), "s"(h.w), "s"(r0.x), "s"(r2.x), "s"(r4.w));
      const bool leaf = (m.x & REC_LEAF) != 0;
      if (skip == i) {
        skip = i + 1;
        if (leaf) {
          const int64_t s = m.x & ~REC_LEAF;
          for (uint32_t j = 0; j < m.y; ++j) {
            const int64_t k = s + j;
            wn += dipole(Ps[3 * k] - qx, Ps[3 * k + 1] - qy, Ps[3 * k + 2] - qz, NAs[3 * k], NAs[3 * k + 1],
                         NAs[3 * k + 2]);
          }
        } else {
          const float lx = h.x - qx, ly = h.y - qy, lz = h.z - qz;
          const float l2 = lx * lx + ly * ly + lz * lz;
          if ((double)sqrtf(l2) > beta * (double)h.w) {
            constexpr int ORDER = NT == 3 ? 0 : NT == 12 ? 1 : 2;
            wn += far_field_poly<ORDER>(lx, ly, lz, l2, r0, r1, r2, r3, r4);
            skip = m.x;
          }
        }
      }
      const bool step = leaf || __ballot(skip == i + 1) != 0;
      i = __builtin_amdgcn_readfirstlane(step ? i + 1 : m.x);
    }
  }
  WN[qi] = wn;
}

// beta <= 0: brute-force dipole sum with a double accumulator (300-310)
__global__ __launch_bounds__(256) void k_cloud_brute(const float* __restrict__ Ps, const float* __restrict__ NAs,
                                                     int64_t npts, const float* __restrict__ Q, int64_t nq,
                                                     float* __restrict__ WN) {
  __shared__ float tile[6 * 256];
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = q < nq;
  const float qx = ok ? Q[3 * q] : 0.f, qy = ok ? Q[3 * q + 1] : 0.f, qz = ok ? Q[3 * q + 2] : 0.f;
  double wn = 0.0;
  for (int64_t b = 0; b < npts; b += 256) {
    const int cnt = (int)((npts - b) < 256 ? (npts - b) : 256);
    __syncthreads();
    if ((int)threadIdx.x < cnt)
      for (int d = 0; d < 3; ++d) {
        tile[6 * threadIdx.x + d] = Ps[3 * (b + threadIdx.x) + d];
        tile[6 * threadIdx.x + 3 + d] = NAs[3 * (b + threadIdx.x) + d];
      }
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const float* t = tile + 6 * j;
      wn += (double)dipole(t[0] - qx, t[1] - qy, t[2] - qz, t[3], t[4], t[5]);
    }
  }
  if (ok) WN[q] = (float)wn;
}

// ---------------- host orchestration ----------------
// CM, R, EC of every cell.  fold_cm: split cells of this library's own
// octree take their centres from the per-level folds (their point lists are
// ranges of the key order, the reference sums in ascending index order);
// otherwise each cell sums its centre over its own list, in list order.
static void cell_moments(gwn_cloud* c, bool fold_cm, const SplitCell* all, const std::vector<uint32_t>& lvl_off,
                         hipStream_t st) {
  const int64_t ncells = c->ncells;
  const int nt = c->nterms;
  c->CM.alloc(3 * (size_t)ncells);
  c->R.alloc(ncells);
  c->EC.alloc((size_t)nt * ncells);
  {
    DevBuf<int32_t> small(ncells), large(ncells);
    DevBuf<uint32_t> ns(2);
    {
      DevBuf<uint32_t> fs(ncells), fl(ncells), ps(ncells), pl(ncells);
      DevBuf<uint8_t> sc(scan_temp_bytes(ncells));
      hipLaunchKernelGGL(k_cell_flags, dim3(div_up(ncells, 256)), dim3(256), 0, st, ncells, c->ccount.p, fs.p, fl.p);
      check_launch("k_cell_flags");
      exclusive_scan_u32(fs.p, ps.p, ncells, ns.p, sc.p, st);
      exclusive_scan_u32(fl.p, pl.p, ncells, ns.p + 1, sc.p, st);
      hipLaunchKernelGGL(k_cell_lists, dim3(div_up(ncells, 256)), dim3(256), 0, st, ncells, fs.p, ps.p, fl.p, pl.p,
                         small.p, large.p);
      check_launch("k_cell_lists");
    }
    uint32_t cnts[2];
    GWN_HIP(hipMemcpyAsync(cnts, ns.p, 8, hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
    hipLaunchKernelGGL(k_cell_empty, dim3(div_up(ncells, 256)), dim3(256), 0, st, ncells, c->ccount.p, nt, c->CM.p,
                       c->R.p, c->EC.p);
    check_launch("k_cell_empty");
    // split-cell centres, level by level.  The work buffers are allocated
    // once for the largest level: a level's cells are disjoint ranges of the
    // points, so it has at most n (cell, point) pairs.
    uint32_t maxm = 0;
    for (size_t d = 0; d + 1 < lvl_off.size(); ++d) maxm = std::max(maxm, lvl_off[d + 1] - lvl_off[d]);
    const bool any_level = fold_cm && maxm > 0;
    const int64_t np = c->n;
    DevBuf<uint32_t> cnt(any_level ? maxm : 0), off(any_level ? maxm : 0), tot(any_level ? 1 : 0);
    DevBuf<uint8_t> sct(any_level ? scan_temp_bytes(maxm) : 0);
    DevBuf<uint64_t> keys(any_level ? np : 0);
    DevBuf<uint32_t> dummy(any_level ? np : 0);
    DevBuf<float4> terms(any_level ? np : 0);
    const size_t tb = any_level ? radix_sort_temp_bytes(np, true) : 0;
    DevBuf<uint8_t> tmp(tb);
    for (size_t d = 0; any_level && d + 1 < lvl_off.size(); ++d) {
      const uint32_t b = lvl_off[d], e = lvl_off[d + 1];
      if (e == b) continue;
      const uint32_t m = e - b;
      hipLaunchKernelGGL(k_split_counts, dim3(div_up(m, 256)), dim3(256), 0, st, all + b, m, cnt.p);
      check_launch("k_split_counts");
      exclusive_scan_u32(cnt.p, off.p, m, tot.p, sct.p, st);
      uint32_t T = 0;
      GWN_HIP(hipMemcpyAsync(&T, tot.p, 4, hipMemcpyDeviceToHost, st));
      GWN_HIP(hipStreamSynchronize(st));
      GWN_CHECK((int64_t)T <= np, GWN_ERR_INVALID, "cloud build: a level holds more pairs than points");
      const unsigned ysplit = m < 4096 ? (unsigned)div_up(4096, m) : 1u;
      hipLaunchKernelGGL(k_lvl_emit, dim3(m, ysplit), dim3(64), 0, st, all + b, off.p, c->perm.p, keys.p);
      check_launch("k_lvl_emit");
      int obits = 1;
      while ((1u << obits) < m) ++obits;
      radix_sort_u64(keys.p, dummy.p, T, 32 + obits, tmp.p, tb, st);
      hipLaunchKernelGGL(k_lvl_terms, dim3(div_up(T, 256)), dim3(256), 0, st, (int64_t)T, keys.p, c->P.p, c->A.p,
                         terms.p);
      check_launch("k_lvl_terms");
      if (m <= 65536) {
        hipLaunchKernelGGL(k_lvl_fold_wave, dim3(m), dim3(64), 0, st, all + b, off.p, terms.p, c->CM.p);
        check_launch("k_lvl_fold_wave");
      } else {
        hipLaunchKernelGGL(k_lvl_fold, dim3(div_up(m, 64)), dim3(64), 0, st, all + b, m, off.p, terms.p, c->CM.p);
        check_launch("k_lvl_fold");
      }
      GWN_HIP(hipStreamSynchronize(st));
    }
    // large cells: centres of the ones not folded above, then their moments
    // in segments of MOMENT_SEGMENT points
    std::vector<int32_t> icell, ij0;
    std::vector<uint32_t> ioff(1, 0);
    if (cnts[1]) {
      hipLaunchKernelGGL(k_cm_fold_wave, dim3(cnts[1]), dim3(64), 0, st, large.p, c->CH.p, fold_cm, c->cstart.p,
                         c->ccount.p, c->P.p, c->A.p, c->perm.p, c->CM.p);
      check_launch("k_cm_fold_wave");
      DevBuf<int32_t> lc(cnts[1]);
      hipLaunchKernelGGL(k_large_counts, dim3(div_up(cnts[1], 256)), dim3(256), 0, st, large.p, cnts[1],
                         c->ccount.p, lc.p);
      check_launch("k_large_counts");
      std::vector<int32_t> hl(cnts[1]), hc(cnts[1]);
      GWN_HIP(hipMemcpyAsync(hl.data(), large.p, 4 * (size_t)cnts[1], hipMemcpyDeviceToHost, st));
      GWN_HIP(hipMemcpyAsync(hc.data(), lc.p, 4 * (size_t)cnts[1], hipMemcpyDeviceToHost, st));
      GWN_HIP(hipStreamSynchronize(st));
      ioff.reserve(cnts[1] + 1);
      for (uint32_t t = 0; t < cnts[1]; ++t) {
        for (int32_t j0 = 0; j0 < hc[t]; j0 += MOMENT_SEGMENT) {
          icell.push_back(hl[t]);
          ij0.push_back(j0);
        }
        ioff.push_back((uint32_t)icell.size());
      }
    }
    const uint32_t nitems = (uint32_t)icell.size();
    DevBuf<int32_t> dcell, dj0;
    DevBuf<uint32_t> doff;
    DevBuf<double> part;
    if (nitems) {
      dcell.alloc(nitems);
      dj0.alloc(nitems);
      doff.alloc(ioff.size());
      part.alloc((size_t)(nt + 1) * nitems);
      GWN_HIP(hipMemcpyAsync(dcell.p, icell.data(), 4 * (size_t)nitems, hipMemcpyHostToDevice, st));
      GWN_HIP(hipMemcpyAsync(dj0.p, ij0.data(), 4 * (size_t)nitems, hipMemcpyHostToDevice, st));
      GWN_HIP(hipMemcpyAsync(doff.p, ioff.data(), 4 * ioff.size(), hipMemcpyHostToDevice, st));
    }
#define GWN_CELL_LAUNCH(NTV)                                                                                        \
  do {                                                                                                              \
    if (cnts[0])                                                                                                    \
      hipLaunchKernelGGL(k_cell_small<NTV>, dim3(div_up(cnts[0], 128)), dim3(128), 0, st, small.p, cnts[0],        \
                         c->CH.p, fold_cm, c->cstart.p, c->ccount.p, c->P.p, c->N.p, c->A.p, c->perm.p, c->CM.p,   \
                         c->R.p, c->EC.p);                                                                          \
    if (nitems) {                                                                                                   \
      hipLaunchKernelGGL(k_cell_large<NTV>, dim3(div_up((int64_t)nitems * 64, 256)), dim3(256), 0, st, dcell.p,    \
                         dj0.p, nitems, c->cstart.p, c->ccount.p, c->P.p, c->N.p, c->A.p, c->perm.p, c->CM.p,      \
                         part.p);                                                                                   \
      hipLaunchKernelGGL(k_cell_combine<NTV>, dim3(div_up(cnts[1], 256)), dim3(256), 0, st, large.p, cnts[1],      \
                         doff.p, part.p, c->R.p, c->EC.p);                                                          \
    }                                                                                                               \
  } while (0)
    if (nt == 3)
      GWN_CELL_LAUNCH(3);
    else if (nt == 12)
      GWN_CELL_LAUNCH(12);
    else
      GWN_CELL_LAUNCH(39);
#undef GWN_CELL_LAUNCH
    check_launch("k_cell_moments");
  }
}

// Rope records over the non-empty non-root cells (split cells given level by
// level in `all`); a leaf record points at leaf_start[cell] in Ps/NAs.
// Walk cost at each cell centre of the cost grid: the records a query there
// visits at beta 2 (the near/far decisions of k_cloud_traverse, without the
// sums) plus the points of the leaves it sums, as min(255, 15 log2(1 + n)).
__global__ __launch_bounds__(256) void k_cloud_cost_grid(const float4* __restrict__ head,
                                                         const uint2* __restrict__ meta, uint32_t nrec, float lx,
                                                         float ly, float lz, float hx, float hy, float hz,
                                                         uint8_t* __restrict__ grid) {
  constexpr int G = kCostGrid;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G * G * G) return;
  const float qx = lx + (c % G + 0.5f) * hx, qy = ly + ((c / G) % G + 0.5f) * hy, qz = lz + (c / (G * G) + 0.5f) * hz;
  uint32_t i = 0, n = 0;
  while (i < nrec) {
    const uint2 m = meta[i];
    ++n;
    if (m.x & REC_LEAF) {
      n += m.y;
      ++i;
      continue;
    }
    const float4 h = head[i];
    const float x = h.x - qx, y = h.y - qy, z = h.z - qz;
    i = sqrtf(x * x + y * y + z * z) > 2.0f * h.w ? m.x : i + 1;
  }
  grid[c] = (uint8_t)fminf(255.0f, 15.0f * log2f(1.0f + (float)n));
}

static void build_cloud_cost_grid(gwn_cloud* c, hipStream_t st) {
  constexpr int G = kCostGrid;
  if (c->root_leaf || c->nrec == 0) return;
  c->cost_grid.alloc((size_t)G * G * G);
  const float* bb = c->bbox;
  hipLaunchKernelGGL(k_cloud_cost_grid, dim3(div_up(G * G * G, 256)), dim3(256), 0, st, c->head.p, c->meta.p,
                     (uint32_t)c->nrec, bb[0], bb[1], bb[2], (bb[3] - bb[0]) / G, (bb[4] - bb[1]) / G,
                     (bb[5] - bb[2]) / G, c->cost_grid.p);
  check_launch("k_cloud_cost_grid");
}

static void rope_records(gwn_cloud* c, const SplitCell* all, const std::vector<uint32_t>& lvl_off,
                         const int64_t* leaf_start, hipStream_t st) {
  const int64_t ncells = c->ncells;
  if (c->root_leaf || lvl_off.size() < 2) {
    c->nrec = 0;
    GWN_HIP(hipStreamSynchronize(st));
    return;
  }
  DevBuf<uint32_t> rc(ncells), pos(ncells);
  // positions start unset (all ones): cells never reached through non-empty
  // parents keep it and get no record
  hipLaunchKernelGGL(k_fill_i32, dim3(div_up(ncells, 256)), dim3(256), 0, st, reinterpret_cast<int32_t*>(pos.p),
                     ncells, -1);
  check_launch("k_fill_i32");
  const int nlev = (int)lvl_off.size() - 1;
  for (int d = nlev - 1; d >= 0; --d) {
    const uint32_t b = lvl_off[d], e = lvl_off[d + 1];
    if (e > b)
      hipLaunchKernelGGL(k_rec_count, dim3(div_up(e - b, 256)), dim3(256), 0, st, all + b, e - b, c->CH.p,
                         c->ccount.p, rc.p);
  }
  check_launch("k_rec_count");
  uint32_t root_rc = 0;
  GWN_HIP(hipMemcpyAsync(&root_rc, rc.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  c->nrec = root_rc - 1;  // the root itself has no record
  if (c->nrec == 0) return;
  for (int d = 0; d < nlev; ++d) {
    const uint32_t b = lvl_off[d], e = lvl_off[d + 1];
    if (e > b)
      hipLaunchKernelGGL(k_rec_pos, dim3(div_up(e - b, 256)), dim3(256), 0, st, all + b, e - b, c->CH.p,
                         c->ccount.p, rc.p, pos.p);
  }
  check_launch("k_rec_pos");
  c->head.alloc(c->nrec);
  c->meta.alloc(c->nrec);
  hipLaunchKernelGGL(k_rec_write, dim3(div_up(ncells, 256)), dim3(256), 0, st, ncells, c->CH.p, c->ccount.p,
                     leaf_start, rc.p, pos.p, c->CM.p, c->R.p, (uint32_t)c->nrec, c->head.p, c->meta.p);
  check_launch("k_rec_write");
  c->poly.alloc(5 * (size_t)c->nrec);
  const unsigned gp = (unsigned)div_up(c->nrec, 256);
  if (c->nterms == 3)
    hipLaunchKernelGGL(k_rec_poly<3>, dim3(gp), dim3(256), 0, st, (uint32_t)c->nrec, c->meta.p, c->EC.p, c->poly.p);
  else if (c->nterms == 12)
    hipLaunchKernelGGL(k_rec_poly<12>, dim3(gp), dim3(256), 0, st, (uint32_t)c->nrec, c->meta.p, c->EC.p, c->poly.p);
  else
    hipLaunchKernelGGL(k_rec_poly<39>, dim3(gp), dim3(256), 0, st, (uint32_t)c->nrec, c->meta.p, c->EC.p, c->poly.p);
  check_launch("k_rec_poly");
  GWN_HIP(hipStreamSynchronize(st));
}

void cloud_build_device(gwn_cloud* c, hipStream_t st, bool moments) {
  const int64_t n = c->n;
  if (n == 0) {
    // a single empty root cell
    c->ncells = 1;
    c->CH.alloc(8);
    c->CN.alloc(3);
    c->Wd.alloc(1);
    c->cstart.alloc(1);
    c->ccount.alloc(1);
    c->CM.alloc(3);
    c->R.alloc(1);
    c->EC.alloc(c->nterms);
    c->perm.alloc(1);
    c->Ps.alloc(3);
    c->NAs.alloc(3);
    const int32_t ch[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    const float zero3[3] = {0, 0, 0};
    const float nan3[3] = {NAN, NAN, NAN};
    const int64_t s0 = 0;
    const int32_t c0 = 0;
    const float z = 0.0f;
    std::vector<float> ec(c->nterms, 0.0f);
    GWN_HIP(hipMemcpyAsync(c->CH.p, ch, sizeof(ch), hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->CN.p, zero3, sizeof(zero3), hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->Wd.p, &z, 4, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->cstart.p, &s0, 8, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->ccount.p, &c0, 4, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->CM.p, nan3, sizeof(nan3), hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->R.p, &z, 4, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->EC.p, ec.data(), 4 * ec.size(), hipMemcpyHostToDevice, st));
    GWN_HIP(hipStreamSynchronize(st));
    c->nrec = 0;
    c->root_leaf = true;
    c->root_start = c->root_count = 0;
    return;
  }
  // root cube: bbox centre and longest side (octree.cpp:139-150), in float
  std::vector<float> hp(3 * (size_t)n);
  GWN_HIP(hipMemcpyAsync(hp.data(), c->P.p, sizeof(float) * hp.size(), hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  float mn[3] = {hp[0], hp[1], hp[2]}, mx[3] = {hp[0], hp[1], hp[2]};
  for (int64_t i = 1; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      const float v = hp[3 * i + d];
      if (v < mn[d]) mn[d] = v;
      if (v > mx[d]) mx[d] = v;
    }
  float ctr[3];
  for (int d = 0; d < 3; ++d) ctr[d] = (mn[d] + mx[d]) / 2.0f;
  float w0 = mx[0] - mn[0];
  for (int d = 1; d < 3; ++d)
    if (mx[d] - mn[d] > w0) w0 = mx[d] - mn[d];
  for (int d = 0; d < 3; ++d) {
    c->bbox[d] = mn[d];
    c->bbox[3 + d] = mx[d];
  }

  // keys and sort: LSD over the two words (low word, then high word), stable
  DevBuf<uint64_t> hi(n), lo(n), lo0(n), key2(n), ls(n);
  DevBuf<uint32_t> idx(n);
  const unsigned g = div_up(n, 256);
  hipLaunchKernelGGL(k_oct_keys, dim3(g), dim3(256), 0, st, c->P.p, n, ctr[0], ctr[1], ctr[2], w0, hi.p, lo.p, idx.p);
  check_launch("k_oct_keys");
  GWN_HIP(hipMemcpyAsync(lo0.p, lo.p, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
  {
    const size_t tb = radix_sort_temp_bytes(n, true);
    DevBuf<uint8_t> tmp(tb);
    radix_sort_u64(lo.p, idx.p, n, 63, tmp.p, tb, st);
    hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, st, hi.p, idx.p, n, key2.p);
    check_launch("k_gather_u64");
    radix_sort_u64(key2.p, idx.p, n, 63, tmp.p, tb, st);
  }
  hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, st, lo0.p, idx.p, n, ls.p);
  check_launch("k_gather_u64");
  const uint64_t* HS = key2.p;  // high words in octree order
  const uint64_t* LS = ls.p;    // low words in octree order

  // points in octree order
  c->Ps.alloc(3 * (size_t)n);
  c->NAs.alloc(3 * (size_t)n);
  c->perm.alloc(n);
  hipLaunchKernelGGL(k_permute_points, dim3(g), dim3(256), 0, st, c->P.p, c->N.p, c->A.p, idx.p, n, c->Ps.p, c->NAs.p,
                     c->perm.p);
  check_launch("k_permute_points");

  // split-cell ranks
  const int64_t nb = n - 1;
  DevBuf<int32_t> lcp(nb > 0 ? nb : 1);
  DevBuf<uint32_t> opens(nb > 0 ? nb : 1), oscan(nb > 0 ? nb : 1), total(1);
  uint32_t S = 0;
  if (nb > 0) {
    hipLaunchKernelGGL(k_oct_lcp, dim3(div_up(nb, 256)), dim3(256), 0, st, HS, LS, n, lcp.p, opens.p);
    check_launch("k_oct_lcp");
    DevBuf<uint8_t> sc(scan_temp_bytes(nb));
    exclusive_scan_u32(opens.p, oscan.p, nb, total.p, sc.p, st);
    GWN_HIP(hipMemcpyAsync(&S, total.p, 4, hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
  }
  const int64_t ncells = 1 + 8 * (int64_t)S;
  c->ncells = ncells;
  c->CH.alloc(8 * (size_t)ncells);
  c->CN.alloc(3 * (size_t)ncells);
  c->Wd.alloc(ncells);
  c->cstart.alloc(ncells);
  c->ccount.alloc(ncells);
  hipLaunchKernelGGL(k_fill_i32, dim3(div_up(8 * ncells, 256)), dim3(256), 0, st, c->CH.p, 8 * ncells, -1);
  check_launch("k_fill_i32");
  {
    const int64_t s0 = 0;
    const int32_t cnt0 = (int32_t)n;
    GWN_HIP(hipMemcpyAsync(c->CN.p, ctr, sizeof(ctr), hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->Wd.p, &w0, 4, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->cstart.p, &s0, 8, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->ccount.p, &cnt0, 4, hipMemcpyHostToDevice, st));
  }
  // level-synchronous split; all split cells kept, level by level
  DevBuf<SplitCell> all(S > 0 ? S : 1);
  DevBuf<uint32_t> cntd(1), errd(1);
  GWN_HIP(hipMemsetAsync(errd.p, 0, 4, st));
  std::vector<uint32_t> lvl_off{0};
  if (S > 0) {
    SplitCell root{0, 0, 0, n, 0, 0};
    GWN_HIP(hipMemcpyAsync(all.p, &root, sizeof(root), hipMemcpyHostToDevice, st));
    lvl_off.push_back(1);
    for (int d = 0; d <= OCT_MAX_SPLIT_DEPTH; ++d) {
      const uint32_t b = lvl_off[d], e = lvl_off[d + 1];
      if (e == b) break;
      GWN_CHECK(e <= S, GWN_ERR_INVALID, "octree split count mismatch");
      GWN_HIP(hipMemsetAsync(cntd.p, 0, 4, st));
      hipLaunchKernelGGL(k_oct_level, dim3(div_up(e - b, 128)), dim3(128), 0, st, all.p + b, e - b, HS, LS, lcp.p,
                         oscan.p, c->CH.p, c->CN.p, c->Wd.p, c->cstart.p, c->ccount.p, all.p + e, cntd.p, S - e,
                         ncells, errd.p);
      check_launch("k_oct_level");
      uint32_t nn = 0, ef = 0;
      GWN_HIP(hipMemcpyAsync(&nn, cntd.p, 4, hipMemcpyDeviceToHost, st));
      GWN_HIP(hipMemcpyAsync(&ef, errd.p, 4, hipMemcpyDeviceToHost, st));
      GWN_HIP(hipStreamSynchronize(st));
      GWN_CHECK(ef == 0, GWN_ERR_INVALID, "octree build: split ranks out of range");
      GWN_CHECK(e + nn <= S, GWN_ERR_INVALID, "octree split count mismatch");
      lvl_off.push_back(e + nn);
    }
    GWN_CHECK(lvl_off.back() == S, GWN_ERR_INVALID, "octree split cells != scan total");
  }

  c->root_leaf = S == 0;
  c->root_start = 0;
  c->root_count = n;
  if (!moments) {  // igl::octree alone
    c->nrec = 0;
    GWN_HIP(hipStreamSynchronize(st));
    return;
  }
  cell_moments(c, true, all.p, lvl_off, st);
  rope_records(c, all.p, lvl_off, c->cstart.p, st);
  build_cloud_cost_grid(c, st);
  GWN_HIP(hipStreamSynchronize(st));
}

// Precompute and rope over a caller's octree: CH (ncells x 8, -1 for
// leaves), point lists as cell_start (ncells + 1) / point_list, optionally
// CM, R, EC already computed (fast_winding_number.h:88-99 takes them as
// inputs).  The cells reachable from cell 0 must form a tree in which a cell
// either has 8 children or none (CH(c,0) == -1), as octree.cpp:7-169 builds
// it and fast_winding_number.cpp:230-290 walks it.
void cloud_build_from_octree_device(gwn_cloud* c, const int32_t* hCH, int64_t ncells, const int64_t* hstart,
                                    const int32_t* hlist, const float* hCM, const float* hR, const float* hEC,
                                    hipStream_t st) {
  const int64_t n = c->n;
  GWN_CHECK(ncells >= 1 && ncells < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad cell count");
  GWN_CHECK(hCH && hstart && (hstart[ncells] == 0 || hlist), GWN_ERR_INVALID, "null octree input");
  GWN_CHECK(hstart[0] == 0, GWN_ERR_INVALID, "cell_start[0] must be 0");
  for (int64_t i = 0; i < ncells; ++i) {
    GWN_CHECK(hstart[i + 1] >= hstart[i], GWN_ERR_INVALID, "cell_start must be non-decreasing");
    GWN_CHECK(hstart[i + 1] - hstart[i] < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "cell list too long");
  }
  const int64_t L = hstart[ncells];
  for (int64_t k = 0; k < L; ++k)
    GWN_CHECK(hlist[k] >= 0 && hlist[k] < n, GWN_ERR_INVALID, "point index out of range in point_indices");
  // breadth-first walk from the root: levels of split cells, reachability
  std::vector<uint8_t> seen((size_t)ncells, 0);
  std::vector<SplitCell> lv;
  std::vector<uint32_t> lvl_off{0};
  std::vector<int32_t> frontier{0}, next, leaves;
  seen[0] = 1;
  for (int depth = 0; !frontier.empty(); ++depth) {
    next.clear();
    for (int32_t cc : frontier) {
      const int32_t* ch = hCH + 8 * (int64_t)cc;
      if (ch[0] == -1) {
        leaves.push_back(cc);
        continue;
      }
      lv.push_back(SplitCell{cc, depth, hstart[cc], hstart[cc + 1], 0u, 0u});
      for (int o = 0; o < 8; ++o) {
        GWN_CHECK(ch[o] >= 0 && ch[o] < ncells, GWN_ERR_INVALID, "CH: a split cell needs 8 valid children");
        GWN_CHECK(!seen[ch[o]], GWN_ERR_INVALID, "CH does not describe a tree (a cell is reached twice)");
        seen[ch[o]] = 1;
        next.push_back(ch[o]);
      }
    }
    if (lv.size() > lvl_off.back()) lvl_off.push_back((uint32_t)lv.size());
    frontier.swap(next);
  }
  c->ncells = ncells;
  c->user_octree = true;
  c->root_leaf = hCH[0] == -1;
  // point data: all points in index order (brute force, beta <= 0), then the
  // lists of the reachable non-empty leaves back to back (rope leaf records)
  std::vector<int64_t> lstart((size_t)ncells, 0);
  std::vector<uint32_t> gather((size_t)n);
  for (int64_t i = 0; i < n; ++i) gather[i] = (uint32_t)i;
  for (int32_t lc : leaves) {
    lstart[lc] = (int64_t)gather.size();
    for (int64_t k = hstart[lc]; k < hstart[lc + 1]; ++k) gather.push_back((uint32_t)hlist[k]);
  }
  GWN_CHECK(gather.size() < (size_t)0x7FFFFFFF, GWN_ERR_INVALID, "leaf lists too long");
  c->root_start = c->root_leaf ? lstart[0] : 0;
  c->root_count = c->root_leaf ? hstart[1] - hstart[0] : 0;
  const int64_t G = (int64_t)gather.size();
  {
    DevBuf<uint32_t> dg(G > 0 ? G : 1);
    DevBuf<int32_t> pout(G > 0 ? G : 1);
    c->Ps.alloc(3 * (size_t)(G > 0 ? G : 1));
    c->NAs.alloc(3 * (size_t)(G > 0 ? G : 1));
    if (G > 0) {
      GWN_HIP(hipMemcpyAsync(dg.p, gather.data(), 4 * (size_t)G, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_permute_points, dim3(div_up(G, 256)), dim3(256), 0, st, c->P.p, c->N.p, c->A.p, dg.p, G,
                         c->Ps.p, c->NAs.p, pout.p);
      check_launch("k_permute_points");
    }
    GWN_HIP(hipStreamSynchronize(st));
  }
  // octree arrays as given; CN and W are not part of the input
  std::vector<int64_t> cs((size_t)ncells);
  std::vector<int32_t> cn((size_t)ncells);
  for (int64_t i = 0; i < ncells; ++i) {
    cs[i] = hstart[i];
    cn[i] = (int32_t)(hstart[i + 1] - hstart[i]);
  }
  c->CH.alloc(8 * (size_t)ncells);
  c->cstart.alloc(ncells);
  c->ccount.alloc(ncells);
  c->perm.alloc(L > 0 ? L : 1);
  GWN_HIP(hipMemcpyAsync(c->CH.p, hCH, 32 * (size_t)ncells, hipMemcpyHostToDevice, st));
  GWN_HIP(hipMemcpyAsync(c->cstart.p, cs.data(), 8 * (size_t)ncells, hipMemcpyHostToDevice, st));
  GWN_HIP(hipMemcpyAsync(c->ccount.p, cn.data(), 4 * (size_t)ncells, hipMemcpyHostToDevice, st));
  if (L > 0) GWN_HIP(hipMemcpyAsync(c->perm.p, hlist, 4 * (size_t)L, hipMemcpyHostToDevice, st));
  // query ordering box: the points' bounding box
  std::vector<float> hp(3 * (size_t)n);
  if (n > 0) GWN_HIP(hipMemcpyAsync(hp.data(), c->P.p, 4 * hp.size(), hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  for (int d = 0; d < 3; ++d) {
    c->bbox[d] = n > 0 ? hp[d] : 0.0f;
    c->bbox[3 + d] = n > 0 ? hp[d] : 0.0f;
  }
  for (int64_t i = 1; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      c->bbox[d] = std::min(c->bbox[d], hp[3 * i + d]);
      c->bbox[3 + d] = std::max(c->bbox[3 + d], hp[3 * i + d]);
    }
  DevBuf<SplitCell> all(lv.size() > 0 ? lv.size() : 1);
  if (!lv.empty())
    GWN_HIP(hipMemcpyAsync(all.p, lv.data(), sizeof(SplitCell) * lv.size(), hipMemcpyHostToDevice, st));
  if (hCM && hR && hEC) {
    c->CM.alloc(3 * (size_t)ncells);
    c->R.alloc(ncells);
    c->EC.alloc((size_t)c->nterms * ncells);
    GWN_HIP(hipMemcpyAsync(c->CM.p, hCM, 12 * (size_t)ncells, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->R.p, hR, 4 * (size_t)ncells, hipMemcpyHostToDevice, st));
    GWN_HIP(hipMemcpyAsync(c->EC.p, hEC, 4 * (size_t)c->nterms * ncells, hipMemcpyHostToDevice, st));
  } else {
    GWN_CHECK(!hCM && !hR && !hEC, GWN_ERR_INVALID, "give all of CM, R, EC or none");
    cell_moments(c, false, all.p, lvl_off, st);
  }
  DevBuf<int64_t> dls(ncells);
  GWN_HIP(hipMemcpyAsync(dls.p, lstart.data(), 8 * (size_t)ncells, hipMemcpyHostToDevice, st));
  rope_records(c, all.p, lvl_off, dls.p, st);
  build_cloud_cost_grid(c, st);
  GWN_HIP(hipStreamSynchronize(st));
}

void cloud_query_device(const gwn_cloud* c, const float* Q, int64_t nq, double beta, float* WN, hipStream_t st) {
  if (nq <= 0) return;
  if (c->n == 0) {
    fill_device<float>(WN, nq, 0.0f, st);
    return;
  }
  const unsigned g = div_up(nq, 256);
  if (!(beta > 0)) {
    ProfileScope prof("cloud_brute", st);
    hipLaunchKernelGGL(k_cloud_brute, dim3(g), dim3(256), 0, st, c->Ps.p, c->NAs.p, c->n, Q, nq, WN);
    check_launch("k_cloud_brute");
    return;
  }
  // scratch of this call: Hilbert keys, the query permutation, sort temp
  const size_t tb = radix_sort_temp_bytes(nq, false);
  Scratch ws(scratch_bytes({4 * (size_t)nq, 4 * (size_t)nq, tb}), st);
  uint32_t* keys = ws.take_n<uint32_t>((size_t)nq);
  uint32_t* perm = ws.take_n<uint32_t>((size_t)nq);
  void* tmp = ws.take(tb);
  hilbert_keys(Q, nq, c->bbox, keys, perm, st);
  radix_sort_u32(keys, perm, nq, 30, tmp, tb, st);
  const bool rl = c->root_leaf;
  // the sort keys are spent: their space holds the block order (g entries)
  // and the block costs (g bytes)
  uint32_t* order = nullptr;
  if (!rl && c->cost_grid.p) {
    order = keys;
    block_order(c->cost_grid.p, c->bbox, Q, perm, nq, order, (uint8_t*)(order + g), st);
  }
  const float* Ps = c->Ps.p + (rl ? 3 * c->root_start : 0);
  const float* NAs = c->NAs.p + (rl ? 3 * c->root_start : 0);
  const int64_t npts = rl ? c->root_count : 0;
  ProfileScope prof("cloud_traverse", st);
  if (c->nterms == 3)
    hipLaunchKernelGGL(k_cloud_traverse<3>, dim3(g), dim3(256), 0, st, c->head.p, c->meta.p, (uint32_t)c->nrec,
                       c->poly.p, Ps, NAs, npts, rl, Q, perm, order, nq, beta, WN);
  else if (c->nterms == 12)
    hipLaunchKernelGGL(k_cloud_traverse<12>, dim3(g), dim3(256), 0, st, c->head.p, c->meta.p, (uint32_t)c->nrec,
                       c->poly.p, Ps, NAs, npts, rl, Q, perm, order, nq, beta, WN);
  else
    hipLaunchKernelGGL(k_cloud_traverse<39>, dim3(g), dim3(256), 0, st, c->head.p, c->meta.p, (uint32_t)c->nrec,
                       c->poly.p, Ps, NAs, npts, rl, Q, perm, order, nq, beta, WN);
  check_launch("k_cloud_traverse");
}

}  // namespace gwn
