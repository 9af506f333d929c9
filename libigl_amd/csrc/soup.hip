// soup.hip -- triangle-soup fast winding number on gfx950.
//
// Pipeline for one mesh (gwn_soup_build):
//   k_tri_setup     per triangle: float box (the reference's K1 loop,
//                   FastWindingNumberForSoups.h:6566-6590) and a packed
//                   a,b,c record for exact evaluation
//   tree            host SAH4 (reference topology) or device LBVH
//   k_parent/k_depth, level buckets
//   k_upward<T>     one launch per tree level, deepest first: per node, the
//                   reference's item()/post() moment formulas -> BoxData,
//                   LocalData for the parent, and the subtree slot count
//   k_positions     one launch per level, root first: first slot of each node
//                   in the rope (skip-pointer) layout
//   k_write_slots   one 96-byte record per non-empty child slot
//
// Query (gwn_soup_query*): k_hilbert_q + radix sort orders queries along a
// Hilbert curve so a wavefront's 64 queries walk similar paths; k_traverse walks
// the rope layout with no stack: a far child slot (|q-P|^2 > maxP2*beta^2 and
// a finite Taylor value) adds its order-0/1/2 term and jumps to its skip
// pointer, a near internal slot steps into its first child slot (the next
// record), a near triangle slot adds the exact Van Oosterom-Strackee angle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "device_math.h"
#include "gwn_internal.h"
#include "ring_walk.h"
#include "soup_moments.h"

namespace gwn {

constexpr uint32_t LINK_LEAF = 0x80000000u;
constexpr double PI_D = 3.1415926535897932384626433832795;

// ---------------------------------------------------------------------------
// setup
__global__ void k_tri_setup(const float* __restrict__ U, const int32_t* __restrict__ F, int64_t nf,
                            float* __restrict__ boxes, float* __restrict__ tri) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int32_t* t = F + 3 * i;
  float p[3][3];
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) p[k][a] = U[3 * (int64_t)t[k] + a];
  for (int a = 0; a < 3; ++a) {
    float lo = p[0][a], hi = p[0][a];
    lo = fmin_h(lo, p[1][a]);
    hi = fmax_h(hi, p[1][a]);
    lo = fmin_h(lo, p[2][a]);
    hi = fmax_h(hi, p[2][a]);
    boxes[6 * i + 2 * a] = lo;
    boxes[6 * i + 2 * a + 1] = hi;
  }
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) tri[9 * i + 3 * k + a] = p[k][a];
}

__global__ void k_parent(const uint32_t* __restrict__ nodes, int64_t nn, int32_t* __restrict__ parent) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  if (i == 0) parent[0] = -1;
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * i + s];
    if (c != NODE_EMPTY && (c & NODE_INTERNAL)) parent[c & ~NODE_INTERNAL] = (int32_t)i;
  }
}

// Depth of every node and the per-depth counts.  Counts are gathered per
// block in LDS first (depths below 64; deeper ones go straight to memory):
// a global atomic per node on the few per-depth counters serialised.
constexpr int kLdsDepths = 64;

__global__ __launch_bounds__(256) void k_depth(const int32_t* __restrict__ parent, int64_t nn,
                                               int32_t* __restrict__ depth, uint32_t* __restrict__ hist,
                                               int32_t* __restrict__ maxdepth) {
  __shared__ uint32_t h[kLdsDepths];
  __shared__ int32_t dmax;
  if (threadIdx.x < kLdsDepths) h[threadIdx.x] = 0;
  if (threadIdx.x == 0) dmax = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) {
    int d = 0;
    int32_t p = parent[i];
    while (p >= 0 && d < 4096) {
      ++d;
      p = parent[p];
    }
    depth[i] = d;
    if (d < kLdsDepths)
      atomicAdd(&h[d], 1u);
    else
      atomicAdd(&hist[d], 1u);
    atomicMax(&dmax, d);
  }
  __syncthreads();
  if (threadIdx.x < kLdsDepths && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
  if (threadIdx.x == 0) atomicMax(maxdepth, dmax);
}

// order[] grouped by depth: each block reserves one range per depth with a
// single atomic and places its nodes inside it (the order within a depth is
// immaterial: the nodes of a level are processed independently)
__global__ __launch_bounds__(256) void k_bucket(const int32_t* __restrict__ depth, int64_t nn,
                                                uint32_t* __restrict__ cursor, int32_t* __restrict__ order) {
  __shared__ uint32_t cnt[kLdsDepths], base[kLdsDepths];
  if (threadIdx.x < kLdsDepths) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = i < nn ? depth[i] : -1;
  uint32_t local = 0;
  if (d >= 0 && d < kLdsDepths) local = atomicAdd(&cnt[d], 1u);
  __syncthreads();
  if (threadIdx.x < kLdsDepths && cnt[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]);
  __syncthreads();
  if (d >= kLdsDepths) order[atomicAdd(&cursor[d], 1u)] = (int32_t)i;
  else if (d >= 0) order[base[d] + local] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// upward pass: one level
template <typename T>
__global__ void k_upward(const int32_t* __restrict__ level_nodes, int64_t count, const uint32_t* __restrict__ nodes,
                         const float* __restrict__ U, const int32_t* __restrict__ F,
                         const float* __restrict__ boxes, int order, Local<T>* __restrict__ local,
                         T* __restrict__ boxdata, uint32_t* __restrict__ nslots_sub) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t node = level_nodes[t];
  Local<T> ch[4];
  int nch = 0;
  uint32_t cnt = 0;
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * (int64_t)node + s];
    if (c == NODE_EMPTY) break;
    if (c & NODE_INTERNAL) {
      const uint32_t ci = c & ~NODE_INTERNAL;
      ch[s] = local[ci];
      cnt += 1 + nslots_sub[ci];
    } else {
      tri_local<T>(U, F, boxes, c, order, ch[s]);
      cnt += 1;
    }
    ++nch;
  }
  T bd[92];
  Local<T> out;
  node_post<T>(ch, nch, order, bd, out);
  local[node] = out;
  for (int k = 0; k < 92; ++k) boxdata[92 * (int64_t)node + k] = bd[k];
  nslots_sub[node] = cnt;
}

// top-down: first slot position of each child node
__global__ void k_positions(const int32_t* __restrict__ level_nodes, int64_t count,
                            const uint32_t* __restrict__ nodes, const uint32_t* __restrict__ nslots_sub,
                            uint32_t* __restrict__ first_slot) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t node = level_nodes[t];
  uint32_t pos = first_slot[node];
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * (int64_t)node + s];
    if (c == NODE_EMPTY) break;
    if (c & NODE_INTERNAL) {
      const uint32_t ci = c & ~NODE_INTERNAL;
      first_slot[ci] = pos + 1;
      pos += 1 + nslots_sub[ci];
    } else {
      pos += 1;
    }
  }
}

// Slot record (6 x float4):
//   0: P.x P.y P.z maxP2          1: -N.x -N.y -N.z link
//   2: A B C Dxy                  3: Dyz Dzx g300 g030
//   4: g003 g210 g201 g120        5: g021 g102 g012 g111
// The far-field terms of the reference (7199-7245), with x = q/|q|:
//   -(x.N)/|q|^2 + (tr Nd - 3 x'Nd x)/|q|^3 + (1.5 (x.s) - 7.5 K(x))/|q|^4,
// s and the cubic K as in the reference (temp0/temp1), are rewritten with
// |x| = 1 as m^2 [ x.(-N) + m ( Qf(x) + m G(x) ) ], m = 1/|q|, where
// Qf(x) = A x^2 + B y^2 + C z^2 + Dxy xy + Dyz yz + Dzx zx with
// A = tr - 3 Ndx ... and Dxy = -3 (Nxy+Nyx) ..., and the cubic
// G(x) = 1.5 (s.x) |x|^2 - 7.5 K(x) with coefficients g_ijk (of x^i y^j z^k).
// The coefficients are formed in FP64 from the reference's FP32 BoxData and
// rounded once.
// link = skip index for an internal child, LINK_LEAF|triangle for an item.
__global__ void k_write_slots(int64_t nn, const uint32_t* __restrict__ nodes, const float* __restrict__ boxdata,
                              const uint32_t* __restrict__ nslots_sub, const uint32_t* __restrict__ first_slot,
                              float4* __restrict__ slots) {
  int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  const float* bd = boxdata + 92 * node;
  uint32_t pos = first_slot[node];
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * node + s];
    if (c == NODE_EMPTY) break;
    uint32_t link, next;
    if (c & NODE_INTERNAL) {
      next = pos + 1 + nslots_sub[c & ~NODE_INTERNAL];
      link = next;
    } else {
      next = pos + 1;
      link = LINK_LEAF | c;
    }
    float4* r = slots + 6 * (int64_t)pos;
    const double nd0 = bd[BD_ND + s], nd1 = bd[BD_ND + 4 + s], nd2 = bd[BD_ND + 8 + s];
    const double tr = nd0 + nd1 + nd2;
    const double k0 = bd[BD_NKD + s], k1 = bd[BD_NKD + 4 + s], k2 = bd[BD_NKD + 8 + s];
    const double cxxy = bd[BD_CXXY + s], cxxz = bd[BD_CXXZ + s], cyyz = bd[BD_CYYZ + s];
    const double cyyx = bd[BD_CYYX + s], czzx = bd[BD_CZZX + s], czzy = bd[BD_CZZY + s];
    const double s0 = 3.0 * k0 + (cyyx + czzx), s1 = 3.0 * k1 + (czzy + cxxy), s2 = 3.0 * k2 + (cxxz + cyyz);
    r[0] = make_float4(bd[BD_P + s], bd[BD_P + 4 + s], bd[BD_P + 8 + s], bd[BD_MAXP2 + s]);
    r[1] = make_float4(-bd[BD_N + s], -bd[BD_N + 4 + s], -bd[BD_N + 8 + s], __uint_as_float(link));
    r[2] = make_float4((float)(tr - 3.0 * nd0), (float)(tr - 3.0 * nd1), (float)(tr - 3.0 * nd2),
                       (float)(-3.0 * (double)bd[BD_NXY + s]));
    r[3] = make_float4((float)(-3.0 * (double)bd[BD_NYZ + s]), (float)(-3.0 * (double)bd[BD_NZX + s]),
                       (float)(1.5 * s0 - 7.5 * k0), (float)(1.5 * s1 - 7.5 * k1));
    r[4] = make_float4((float)(1.5 * s2 - 7.5 * k2), (float)(1.5 * s1 - 7.5 * cxxy), (float)(1.5 * s2 - 7.5 * cxxz),
                       (float)(1.5 * s0 - 7.5 * cyyx));
    r[5] = make_float4((float)(1.5 * s2 - 7.5 * cyyz), (float)(1.5 * s0 - 7.5 * czzx), (float)(1.5 * s1 - 7.5 * czzy),
                       (float)(-7.5 * (double)bd[BD_SP + s]));
    pos = next;
  }
}

// FP64 slot record: 24 doubles in the same field order (link kept in the FP32 record).
__global__ void k_write_slots64(int64_t nn, const uint32_t* __restrict__ nodes, const double* __restrict__ bd64,
                                const uint32_t* __restrict__ first_slot, const uint32_t* __restrict__ nslots_sub,
                                double* __restrict__ slots64) {
  int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  const double* bd = bd64 + 92 * node;
  uint32_t pos = first_slot[node];
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * node + s];
    if (c == NODE_EMPTY) break;
    double* r = slots64 + 24 * (int64_t)pos;
    r[0] = bd[BD_P + s];
    r[1] = bd[BD_P + 4 + s];
    r[2] = bd[BD_P + 8 + s];
    r[3] = bd[BD_MAXP2 + s];
    r[4] = bd[BD_N + s];
    r[5] = bd[BD_N + 4 + s];
    r[6] = bd[BD_N + 8 + s];
    r[7] = 0.0;
    r[8] = bd[BD_ND + s];
    r[9] = bd[BD_ND + 4 + s];
    r[10] = bd[BD_ND + 8 + s];
    r[11] = bd[BD_NXY + s];
    r[12] = bd[BD_NYZ + s];
    r[13] = bd[BD_NZX + s];
    r[14] = bd[BD_NKD + s];
    r[15] = bd[BD_NKD + 4 + s];
    r[16] = bd[BD_NKD + 8 + s];
    r[17] = bd[BD_SP + s];
    r[18] = bd[BD_CXXY + s];
    r[19] = bd[BD_CXXZ + s];
    r[20] = bd[BD_CYYZ + s];
    r[21] = bd[BD_CYYX + s];
    r[22] = bd[BD_CZZX + s];
    r[23] = bd[BD_CZZY + s];
    pos += (c & NODE_INTERNAL) ? 1 + nslots_sub[c & ~NODE_INTERNAL] : 1;
  }
}

// ---------------------------------------------------------------------------
// closest-point support: a bounding box per slot record, from the FP64
// vertices and rounded outward to float, so a box distance computed from it
// never exceeds the distance to anything inside.
__device__ __forceinline__ float round_down_f(double x) {
  float f = (float)x;
  return ((double)f > x) ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float round_up_f(double x) {
  float f = (float)x;
  return ((double)f < x) ? nextafterf(f, INFINITY) : f;
}
__device__ __forceinline__ void tri_box64(const double* __restrict__ V, const int32_t* __restrict__ F, uint32_t t,
                                          float lo[3], float hi[3]) {
  for (int a = 0; a < 3; ++a) {
    double l = V[3 * (int64_t)F[3 * (int64_t)t] + a], h = l;
    for (int k = 1; k < 3; ++k) {
      const double v = V[3 * (int64_t)F[3 * (int64_t)t + k] + a];
      l = v < l ? v : l;
      h = v > h ? v : h;
    }
    lo[a] = round_down_f(l);
    hi[a] = round_up_f(h);
  }
}

__global__ void k_node_boxes(const int32_t* __restrict__ level_nodes, int64_t count, const uint32_t* __restrict__ nodes,
                             const double* __restrict__ V, const int32_t* __restrict__ F, float* __restrict__ nbox) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t node = level_nodes[t];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * (int64_t)node + s];
    if (c == NODE_EMPTY) break;
    float cl[3], chh[3];
    if (c & NODE_INTERNAL) {
      const float* b = nbox + 6 * (int64_t)(c & ~NODE_INTERNAL);
      for (int a = 0; a < 3; ++a) {
        cl[a] = b[a];
        chh[a] = b[3 + a];
      }
    } else {
      tri_box64(V, F, c, cl, chh);
    }
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], cl[a]);
      hi[a] = fmaxf(hi[a], chh[a]);
    }
  }
  for (int a = 0; a < 3; ++a) {
    nbox[6 * (int64_t)node + a] = lo[a];
    nbox[6 * (int64_t)node + 3 + a] = hi[a];
  }
}

// first slot of each child node from the parent's first slot and the links
__global__ void k_positions_rope(const int32_t* __restrict__ level_nodes, int64_t count,
                                 const uint32_t* __restrict__ nodes, const float4* __restrict__ slots,
                                 uint32_t* __restrict__ first_slot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t node = level_nodes[t];
  uint32_t pos = first_slot[node];
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * (int64_t)node + s];
    if (c == NODE_EMPTY) break;
    if (c & NODE_INTERNAL) {
      first_slot[c & ~NODE_INTERNAL] = pos + 1;
      pos = __float_as_uint(slots[6 * (int64_t)pos + 1].w);
    } else {
      pos += 1;
    }
  }
}

// slot box record (2 x float4): [lo.xyz, link] [hi.xyz, 0], in rope order
__global__ void k_write_dbox(int64_t nn, const uint32_t* __restrict__ nodes, const float* __restrict__ nbox,
                             const double* __restrict__ V, const int32_t* __restrict__ F,
                             const uint32_t* __restrict__ first_slot, const float4* __restrict__ slots,
                             float4* __restrict__ dbox) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  uint32_t pos = first_slot[node];
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * node + s];
    if (c == NODE_EMPTY) break;
    float lo[3], hi[3];
    if (c & NODE_INTERNAL) {
      const float* b = nbox + 6 * (int64_t)(c & ~NODE_INTERNAL);
      for (int a = 0; a < 3; ++a) {
        lo[a] = b[a];
        hi[a] = b[3 + a];
      }
    } else {
      tri_box64(V, F, c, lo, hi);
    }
    const float link = slots[6 * (int64_t)pos + 1].w;
    dbox[2 * (int64_t)pos] = make_float4(lo[0], lo[1], lo[2], link);
    dbox[2 * (int64_t)pos + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    const uint32_t l = __float_as_uint(link);
    pos = (l & LINK_LEAF) ? pos + 1 : l;
  }
}

// ---------------------------------------------------------------------------
// exact triangle solid angle: UTsignedSolidAngleTri (6070-6110)
template <typename T>
__device__ __forceinline__ T tri_solid_angle(T ax, T ay, T az, T bx, T by, T bz, T cx, T cy, T cz, T qx, T qy,
                                             T qz) {
  T a0 = ax - qx, a1 = ay - qy, a2 = az - qz;
  T b0 = bx - qx, b1 = by - qy, b2 = bz - qz;
  T c0 = cx - qx, c1 = cy - qy, c2 = cz - qz;
  const T al = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
  const T bl = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
  const T cl = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  if (al == T(0) || bl == T(0) || cl == T(0)) return T(0);
  const T ra = T(1) / al, rb = T(1) / bl, rc = T(1) / cl;
  a0 *= ra; a1 *= ra; a2 *= ra;
  b0 *= rb; b1 *= rb; b2 *= rb;
  c0 *= rc; c1 *= rc; c2 *= rc;
  const T u0 = b0 - a0, u1 = b1 - a1, u2 = b2 - a2;
  const T w0 = c0 - a0, w1 = c1 - a1, w2 = c2 - a2;
  const T num = a0 * (u1 * w2 - u2 * w1) + a1 * (u2 * w0 - u0 * w2) + a2 * (u0 * w1 - u1 * w0);
  if (num == T(0)) return T(0);
  const T den = T(1) + (a0 * b0 + a1 * b1 + a2 * b2) + (a0 * c0 + a1 * c1 + a2 * c2) + (b0 * c0 + b1 * c1 + b2 * c2);
  return T(2) * atan2(num, den);
}

// ---- FP32 fast path of the traversal (atan2_fast: device_math.h) ----
// UTsignedSolidAngleTri (6070-6110) with hardware reciprocal square roots for
// the normalisation; same zero-length and zero-numerator rules.
__device__ __forceinline__ float tri_solid_angle_fast(float ax, float ay, float az, float bx, float by, float bz,
                                                      float cx, float cy, float cz, float qx, float qy, float qz) {
  float a0 = ax - qx, a1 = ay - qy, a2 = az - qz;
  float b0 = bx - qx, b1 = by - qy, b2 = bz - qz;
  float c0 = cx - qx, c1 = cy - qy, c2 = cz - qz;
  const float la = fmaf(a2, a2, fmaf(a1, a1, a0 * a0));
  const float lb = fmaf(b2, b2, fmaf(b1, b1, b0 * b0));
  const float lc = fmaf(c2, c2, fmaf(c1, c1, c0 * c0));
  if (la == 0.0f || lb == 0.0f || lc == 0.0f) return 0.0f;
  if (fminf(la, fminf(lb, lc)) < 1e-30f)  // keep rsq away from denormals
    return tri_solid_angle<float>(ax, ay, az, bx, by, bz, cx, cy, cz, qx, qy, qz);
  const float ra = __builtin_amdgcn_rsqf(la), rb = __builtin_amdgcn_rsqf(lb), rc = __builtin_amdgcn_rsqf(lc);
  a0 *= ra; a1 *= ra; a2 *= ra;
  b0 *= rb; b1 *= rb; b2 *= rb;
  c0 *= rc; c1 *= rc; c2 *= rc;
  const float u0 = b0 - a0, u1 = b1 - a1, u2 = b2 - a2;
  const float w0 = c0 - a0, w1 = c1 - a1, w2 = c2 - a2;
  const float k0 = fmaf(u1, w2, -(u2 * w1)), k1 = fmaf(u2, w0, -(u0 * w2)), k2 = fmaf(u0, w1, -(u1 * w0));
  const float num = fmaf(a2, k2, fmaf(a1, k1, a0 * k0));
  if (num == 0.0f) return 0.0f;
  const float ab = fmaf(a2, b2, fmaf(a1, b1, a0 * b0));
  const float ac = fmaf(a2, c2, fmaf(a1, c1, a0 * c0));
  const float bc = fmaf(b2, c2, fmaf(b1, c1, b0 * c0));
  const float den = ((1.0f + ab) + ac) + bc;
  return 2.0f * atan2_fast(num, den);
}

// Far-field Taylor value of one child slot (SolidAngleFunctors::pre,
// 7199-7245) from the reference's moment layout (FP64 records), for
// q = query - P already formed and l2 = |q|^2.
template <int ORDER, typename T, typename V4>
__device__ __forceinline__ T far_field(T x, T y, T z, T l2, const V4& b, const V4& c, const V4& d, const V4& e,
                                       const V4& f) {
#pragma clang fp contract(fast)
  const T m2 = T(1) / l2;
  const T m1 = sqrt(m2);
  x *= m1;
  y *= m1;
  z *= m1;
  T o = -m2 * (x * b.x + y * b.y + z * b.z);
  if (ORDER >= 1) {
    const T x2 = x * x, y2 = y * y, z2 = z * z;
    const T m3 = m2 * m1;
    o += m3 * (c.x + c.y + c.z - T(3) * ((x2 * c.x + y2 * c.y + z2 * c.z) + x * y * c.w + x * z * d.y + y * z * d.x));
    if (ORDER >= 2) {
      const T m4 = m2 * m2;
      // temp0 / temp1 of the reference
      const T t00 = f.y + f.z, t01 = f.w + e.z, t02 = e.w + f.x;
      const T t10 = y * e.z + z * e.w, t11 = z * f.x + x * f.y, t12 = x * f.z + y * f.w;
      const T s0 = T(3) * d.z + t00, s1 = T(3) * d.w + t01, s2 = T(3) * e.x + t02;
      o += m4 * (T(1.5) * (x * s0 + y * s1 + z * s2) -
                 T(7.5) * ((x2 * x * d.z + y2 * y * d.w + z2 * z * e.x) + x * y * z * e.y + (x2 * t10 + y2 * t11 + z2 * t12)));
    }
  }
  return o;
}

struct D4 {
  double x, y, z, w;
};

// ---------------------------------------------------------------------------
// traversal: FP32 (reference arithmetic) and mixed (FP32 far field and
// running sum; FP64 solid angles and sum of the triangles near the query)
//
// Wave-coherent rope walk.  The slot records are in depth-first ("rope")
// order: a far child adds its Taylor value and skips to `link` (the record
// after its subtree), a near internal child is entered (next record), an item
// is evaluated exactly when near.  Each lane keeps `skip`, the next record it
// needs; the wavefront walks one shared index i and a lane takes part in a
// record exactly when i == skip.  (k_traverse tests skip <= i: a lane that
// enters a child or evaluates an item leaves skip where it was, and only a
// lane that takes a far field moves it past the subtree.)  The wave steps to
// i+1 when some lane needs it (a near child, or an item) and otherwise to
// link(i): every lane that is not taking part needs a record at or beyond
// link(i), because the subtrees of the rope are nested intervals.  So each lane visits exactly the records
// of its own walk, in the same order, and sums the same terms in the same
// order as an independent per-lane walk; but every record (and triangle) a
// wave reads is the same for all its lanes, so it is fetched once per wave
// through the scalar cache instead of 64 times through the vector memory path.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Mixed form: a triangle slot that may lie within near_d of the query
// (|q - P| < sqrt(maxP2) + near_d, P the slot's centre and maxP2 its squared
// radius) has its solid angle computed in FP64 (UTsignedSolidAngleTri,
// 6070-6110, in double) and summed in an FP64 accumulator of its own; the far
// field and every other exact term go to the FP32 running sum.
__device__ __forceinline__ bool near_triangle(float maxp2, float l2, float near_d) {
  const float rr = sqrtf(maxp2) + near_d;
  return l2 < rr * rr;
}

// Near-triangle FP64 terms of the mixed form are deferred: a lane notes each
// such triangle in its own column of an LDS list and evaluates the list, in
// the order noted, after the walk.  In the wave-coherent walk this turns FP64
// work that one lane of a wavefront needs at a record (which every lane would
// sit through) into work all lanes do together at the end, and keeps the FP64
// code and its registers out of the walk.  The terms and the order of their
// FP64 additions are those of evaluating each where it is met.  A query that
// meets more than kDefer near triangles gets W = NaN, which sends it to the
// FP64 re-evaluation pass with the ambiguous queries.
constexpr int kDefer = 16;
constexpr int kTravThreads = 256;

struct NearList {
  uint32_t* lds;  // column threadIdx.x of a kDefer x kTravThreads array
  uint32_t n;     // kDefer + 1 after an overflow
};

// Exact term of a near triangle slot: added to sum (FP32) or, in the mixed
// form for a triangle near the query, noted for the FP64 sum.
template <bool MIXED>
__device__ __forceinline__ void exact_term(const float* __restrict__ tri, uint32_t t, float maxp2, float l2, float qx,
                                           float qy, float qz, float near_d, float& sum, NearList& L) {
  if (MIXED && near_triangle(maxp2, l2, near_d)) {
    if (L.n < (uint32_t)kDefer) L.lds[L.n * kTravThreads] = t;
    L.n = L.n < (uint32_t)kDefer ? L.n + 1 : (uint32_t)kDefer + 1;
  } else {
    const float* p = tri + 9 * (size_t)t;
    sum = __fadd_rn(sum, tri_solid_angle_fast(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], qx, qy, qz));
  }
}

// W of a walk: the FP32 sum, plus (mixed form) the noted near triangles'
// FP64 solid angles summed in order; NaN after an overflow.
template <bool MIXED>
__device__ __forceinline__ double finish_walk(float sum, const NearList& L, const float* __restrict__ tri, float qx,
                                              float qy, float qz) {
  if (!MIXED) return (double)sum / (4.0 * PI_D);
  if (L.n > (uint32_t)kDefer) return __longlong_as_double(0x7FF8000000000000ll);
  double sum_near = 0.0;
  for (uint32_t k = 0; k < L.n; ++k) {
    const float* p = tri + 9 * (size_t)L.lds[k * kTravThreads];
    sum_near = __dadd_rn(sum_near,
                         tri_solid_angle<double>(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], qx, qy, qz));
  }
  return ((double)sum + sum_near) / (4.0 * PI_D);
}

// Both walks are held to 8 wavefronts per SIMD (64 VGPRs): the walk loop fits
// in that, and only the mixed form's end-of-walk FP64 code would take more
// (94 VGPRs, 5 wavefronts); it keeps some of its constants in scratch instead.
// Per-lane rope walk from record j to the end: each lane follows its own skip
// pointers and gathers its own records.  The wave-coherent walk hands over to
// it where its wavefront's lanes have stopped sharing records; records, terms
// and their order are the same in both walks.
template <int ORDER, bool MIXED>
__device__ __forceinline__ void lane_walk(const float4* __restrict__ slots, uint32_t nslots,
                                          const float* __restrict__ tri, uint32_t j, float qx, float qy, float qz,
                                          float acc2, float near_d, float& sum, NearList& L) {
  while (j < nslots) {
    const float4* r = slots + 6 * (size_t)j;
    const float4 a = r[0], b = r[1];
    const uint32_t link = __float_as_uint(b.w);
    const float x = qx - a.x, y = qy - a.y, z = qz - a.z;
    const float l2 = __fadd_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)), __fmul_rn(z, z));
    const bool leaf = (link & LINK_LEAF) != 0;
    if (!(l2 <= __fmul_rn(a.w, acc2))) {
      float4 c, d, e, f;
      if (ORDER >= 1) {
        c = r[2];
        d = r[3];
      }
      if (ORDER >= 2) {
        e = r[4];
        f = r[5];
      }
      const float o = far_field_poly<ORDER>(x, y, z, l2, b, c, d, e, f);
      if (isfinite(o)) {
        sum = __fadd_rn(sum, o);
        j = leaf ? j + 1 : link;
        continue;
      }
    }
    if (leaf) exact_term<MIXED>(tri, link & ~LINK_LEAF, a.w, l2, qx, qy, qz, near_d, sum, L);
    ++j;
  }
}

// A kernel argument copied into a VGPR: a VALU instruction reads at most one
// SGPR, so the product of the slot radius (an SGPR in the wave-coherent walk)
// and accuracy_scale^2 would otherwise need a move at every step.
__device__ __forceinline__ float in_vgpr(float x) {
  float v;
  asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
  return v;
}

// HANDOVER: a wavefront whose lanes have stopped sharing records (fewer than
// 16 lanes per step on average over a window of 256 steps: the union of their
// paths exceeds the longest by far, as on fans of thin triangles) hands over
// to the per-lane walk, each lane from the record it needs next.  Results are
// unchanged (same records, terms and order); the check costs about a tenth of
// the walk's time on coherent batches, so only latency-bound callers (the
// serving engine, on large trees) take it.
template <int ORDER, bool MIXED, bool HANDOVER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_traverse(
    const float4* __restrict__ slots, uint32_t nslots, const float* __restrict__ tri, const float* __restrict__ Q,
    const uint32_t* __restrict__ perm, const uint32_t* __restrict__ order, int64_t nq, float acc2_arg, float near_d,
    double* __restrict__ W) {
  const int64_t t = (int64_t)(order ? order[xcd_block()] : xcd_block()) * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const float acc2 = in_vgpr(acc2_arg);
  const int64_t qi = perm ? (int64_t)perm[t] : t;
  const float qx = Q[3 * qi], qy = Q[3 * qi + 1], qz = Q[3 * qi + 2];
  float sum = 0.0f;
  __shared__ uint32_t near_lds[MIXED ? kDefer * kTravThreads : 1];  // mixed form: noted near triangles
  NearList L{near_lds + threadIdx.x, 0};
  uint32_t skip = 0;
  uint32_t i = 0;
  uint32_t steps = 0, taken = 0;  // wave-uniform: steps in this window, lanes taking part summed over them
  while (i < nslots) {
    if (HANDOVER) taken += (uint32_t)__popcll(__ballot(skip <= i));  // lanes taking part in record i
    const float4* r = slots + 6 * (size_t)i;
    const float4 a = r[0], b = r[1];
    float4 c, d, e, f;
    if (ORDER >= 1) {
      c = r[2];
      d = r[3];
    }
    if (ORDER >= 2) {
      e = r[4];
      f = r[5];
    }
    // the whole record is requested together (one scalar-cache round trip)
    // rather than where each part is first used
    if (ORDER >= 2) asm volatile("" ::"s"(a.x), "s"(b.w), "s"(c.x), "s"(e.x), "s"(f.w));
    else if (ORDER >= 1) asm volatile("" ::"s"(a.x), "s"(b.w), "s"(c.x), "s"(d.w));
    else asm volatile("" ::"s"(a.x), "s"(b.w));
    const uint32_t link = __float_as_uint(b.w);
    const bool leaf = (link & LINK_LEAF) != 0;
    // the record after this one's subtree, picked in scalar code so that a
    // lane taking the far field only copies it
    const uint32_t past = leaf ? i + 1 : link;
    // skip >= i holds for every lane (nested intervals), so skip <= i means
    // skip == i; a lane that descends or evaluates an item leaves skip at i,
    // which is <= i + 1, the record it needs next
    if (skip <= i) {
      const float x = qx - a.x, y = qy - a.y, z = qz - a.z;
      // |q-P|^2 rounded as the reference rounds it (no fused multiply-add),
      // so every near/far decision is the reference's
      const float l2 = __fadd_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)), __fmul_rn(z, z));
      bool exact = leaf;
      if (!(l2 <= __fmul_rn(a.w, acc2))) {
        const float o = far_field_poly<ORDER>(x, y, z, l2, b, c, d, e, f);
        if (isfinite(o)) {  // otherwise the reference descends / evaluates the item
          sum = __fadd_rn(sum, o);
          exact = false;
          skip = past;
        }
      }
      if (exact) exact_term<MIXED>(tri, link & ~LINK_LEAF, a.w, l2, qx, qy, qz, near_d, sum, L);
    }
    const bool step = leaf || __ballot(skip <= i + 1) != 0;
    i = wave_uniform(step ? i + 1 : link);
    if (HANDOVER && ++steps == 256) {
      if (taken < 256u * 16u) break;
      steps = taken = 0;
    }
  }
  // after a hand-over each lane continues from the record it needs next
  if (HANDOVER && i < nslots)
    lane_walk<ORDER, MIXED>(slots, nslots, tri, skip > i ? skip : i, qx, qy, qz, acc2, near_d, sum, L);
  W[qi] = finish_walk<MIXED>(sum, L, tri, qx, qy, qz);
}

// all-FP64 form: FP64 moments, far field, atan2 and sum
template <int ORDER>
// Eight wavefronts per SIMD (64 VGPRs, some spilled): its 92 VGPRs allowed five, and on the
// 1,310,720-triangle workload (10M queries) the walk took 59.1 ms at five, 56.7 at six, 54.3 at
// seven and 52.8 at eight.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_traverse64(const float4* __restrict__ slots, const double* __restrict__ s64,
                                                    uint32_t nslots, const float* __restrict__ tri,
                                                    const float* __restrict__ Q, const uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ order, int64_t nq,
                                                    const uint32_t* __restrict__ dcount, double acc2,
                                                    double* __restrict__ W) {
  // the same wave-coherent rope walk as k_traverse, over the FP64 records;
  // dcount (when given) holds the number of queries, read on the device;
  // order (when given) is the block order
  const int64_t t = (int64_t)(order ? order[xcd_block()] : xcd_block()) * blockDim.x + threadIdx.x;
  if (t >= (dcount ? (int64_t)*dcount : nq)) return;
  const int64_t qi = perm ? (int64_t)perm[t] : t;
  const double qx = Q[3 * qi], qy = Q[3 * qi + 1], qz = Q[3 * qi + 2];
  double sum = 0.0;
  uint32_t skip = 0;
  uint32_t i = 0;
  while (i < nslots) {
    const double* r = s64 + 24 * (size_t)i;
    const uint32_t link = __float_as_uint(slots[6 * (size_t)i + 1].w);
    const double r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    // the link and the record's point and radius in one scalar-cache round
    // trip (the compiler would otherwise wait for the link first)
    asm volatile("" ::"s"(link), "s"(r0), "s"(r3));
    const bool leaf = (link & LINK_LEAF) != 0;
    if (skip == i) {
      const double x = qx - r0, y = qy - r1, z = qz - r2;
      const double l2 = x * x + y * y + z * z;
      bool exact = leaf;
      skip = i + 1;
      if (!(l2 <= r3 * acc2)) {
        const D4 b{r[4], r[5], r[6], 0.0}, c{r[8], r[9], r[10], r[11]}, d{r[12], r[13], r[14], r[15]},
            e{r[16], r[17], r[18], r[19]}, f{r[20], r[21], r[22], r[23]};
        const double o = far_field<ORDER, double, D4>(x, y, z, l2, b, c, d, e, f);
        if (isfinite(o)) {
          sum = __dadd_rn(sum, o);
          exact = false;
          if (!leaf) skip = link;
        }
      }
      if (exact) {
        const float* p = tri + 9 * (size_t)(link & ~LINK_LEAF);
        sum = __dadd_rn(sum, tri_solid_angle<double>(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], qx, qy, qz));
      }
    }
    const bool step = leaf || __ballot(skip == i + 1) != 0;
    i = wave_uniform(step ? i + 1 : link);
  }
  W[qi] = sum / (4.0 * PI_D);
}

// ---- FP64 re-evaluation of a few scattered queries: one wavefront per query ----
// A sequential walk is a chain of dependent record fetches as long as the
// query's path (thousands on a large mesh): a pass over a few hundred queries
// runs a handful of wavefronts for milliseconds.  Here the 64 lanes of a
// wavefront share one query's walk (ring_walk, ring_walk.h).  The records
// visited and the terms are those of the sequential walk (each record's
// decision depends on the record and the query only); their FP64 sum is taken
// per lane and then across lanes in a fixed order, so the result is
// deterministic and differs from the sequential sum only by rounding.

// FP64 term of record i for query q: adds a far-field or exact term to part
// and returns true, or returns false when the walk descends into the record.
template <int ORDER>
__device__ __forceinline__ bool record_term64(const double* __restrict__ s64, const float* __restrict__ tri,
                                              uint32_t i, uint32_t link, double qx, double qy, double qz, double acc2,
                                              double& part) {
  const double* r = s64 + 24 * (size_t)i;
  const bool leaf = (link & LINK_LEAF) != 0;
  const double x = qx - r[0], y = qy - r[1], z = qz - r[2];
  const double l2 = x * x + y * y + z * z;
  if (!(l2 <= r[3] * acc2)) {
    const D4 b{r[4], r[5], r[6], 0.0}, c{r[8], r[9], r[10], r[11]}, d{r[12], r[13], r[14], r[15]},
        e{r[16], r[17], r[18], r[19]}, f{r[20], r[21], r[22], r[23]};
    const double o = far_field<ORDER, double, D4>(x, y, z, l2, b, c, d, e, f);
    if (isfinite(o)) {
      part = __dadd_rn(part, o);
      return true;
    }
  }
  if (!leaf) return false;
  const float* p = tri + 9 * (size_t)(link & ~LINK_LEAF);
  part = __dadd_rn(part, tri_solid_angle<double>(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], qx, qy, qz));
  return true;
}


// FP64 ring walk of query q: W = sum / 4 pi, the same on every lane
template <int ORDER>
__device__ __forceinline__ double ring_walk64(uint2* __restrict__ ring, int lane, const float4* __restrict__ slots,
                                              const double* __restrict__ s64, uint32_t nslots,
                                              const float* __restrict__ tri, double qx, double qy, double qz,
                                              double acc2) {
  double part = 0.0;
  ring_walk(ring, lane, nslots, [&](uint32_t i, uint32_t& link) {
    link = __float_as_uint(slots[6 * (size_t)i + 1].w);
    return record_term64<ORDER>(s64, tri, i, link, qx, qy, qz, acc2, part);
  });
  return wave_sum(part) / (4.0 * PI_D);
}

// list[k] (k itself when list is null) for k < count (count = *dcount when
// given, else n): 4 queries per 256-thread block, one per wavefront,
// grid-strided.
template <int ORDER>
__global__ __launch_bounds__(256) void k_reeval64(const float4* __restrict__ slots, const double* __restrict__ s64,
                                                  uint32_t nslots, const float* __restrict__ tri,
                                                  const float* __restrict__ Q, const uint32_t* __restrict__ list,
                                                  int64_t n, const uint32_t* __restrict__ dcount, double acc2,
                                                  double* __restrict__ W) {
  __shared__ uint2 ring_all[4][kReevalRing];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t count = dcount ? (int64_t)*dcount : n;
  for (int64_t k = (int64_t)blockIdx.x * 4 + wv; k < count; k += (int64_t)gridDim.x * 4) {
    const int64_t qi = list ? (int64_t)list[k] : k;
    const double w = ring_walk64<ORDER>(ring_all[wv], lane, slots, s64, nslots, tri, Q[3 * qi], Q[3 * qi + 1],
                                        Q[3 * qi + 2], acc2);
    if (lane == 0) W[qi] = w;
  }
}

// ---- small batches: one wavefront per query for the whole query ----
// A batch of a few thousand queries fills few wavefronts with the lane walk,
// and each lane's sequential walk is a chain of dependent record fetches as
// long as its path: at accuracy_scale 4.9 on a 434k-triangle mesh a batch of
// 1,000 queries takes 1.4 ms however few queries it has.  Here each query gets
// a wavefront and the ring walk above, so a query takes depth-many rounds.
// The records visited and each term are those of the sequential walk (FP32
// decisions and terms with the reference's rounding, FP64 near triangles in
// the mixed form); the sum is taken per lane and then across lanes in a fixed
// order, so results are deterministic and differ from the sequential walks'
// running sums only by rounding.  In the mixed form k_reeval64_band then
// walks the queries whose W falls in the ambiguous band again in FP64, so the
// batch needs no collection pass and no count.
// Mixed form: a lane notes the near triangles it meets (up to kRingDefer, in
// LDS) and sums their FP64 solid angles after the walk, keeping the FP64 code
// out of the walk's registers; a query with more than kRingDefer on one lane
// gets W = NaN, which k_reeval64_band walks again in FP64.
constexpr int kRingDefer = 4;

// FP32 and mixed forms (the FP64 form is k_reeval64 over every query).  Held
// to 8 wavefronts per SIMD: the walk fits in 64 VGPRs; the mixed form's FP64
// tail would take more and keeps some of its values in scratch instead.
template <int ORDER, bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_walk_ring(
    const float4* __restrict__ slots, uint32_t nslots, const float* __restrict__ tri, const float* __restrict__ Q,
    int64_t nq, float acc2f, float near_d, double* __restrict__ W) {
  __shared__ uint2 ring_all[4][kReevalRing];
  __shared__ uint32_t near_all[MIXED ? 4 * kRingDefer * 64 : 1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint2* ring = ring_all[wv];
  uint32_t* near = near_all + (MIXED ? wv * kRingDefer * 64 + lane : 0);  // this lane's column
  for (int64_t qi = (int64_t)blockIdx.x * 4 + wv; qi < nq; qi += (int64_t)gridDim.x * 4) {
    const float qx = Q[3 * qi], qy = Q[3 * qi + 1], qz = Q[3 * qi + 2];
    double w;
    {
      float p32 = 0.0f;
      uint32_t nn = 0;  // near triangles noted (kRingDefer + 1 after an overflow)
      ring_walk(ring, lane, nslots, [&](uint32_t i, uint32_t& link) {
        const float4* r = slots + 6 * (size_t)i;
        // the whole record in one round trip (loading the far-field part only
        // where it is used measured no faster)
        const float4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4], f = r[5];
        link = __float_as_uint(b.w);
        const bool leaf = (link & LINK_LEAF) != 0;
        const float x = qx - a.x, y = qy - a.y, z = qz - a.z;
        // rounded as the reference rounds it (no fused multiply-add)
        const float l2 = __fadd_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)), __fmul_rn(z, z));
        if (!(l2 <= __fmul_rn(a.w, acc2f))) {
          const float o = far_field_poly<ORDER>(x, y, z, l2, b, c, d, e, f);
          if (isfinite(o)) {
            p32 = __fadd_rn(p32, o);
            return true;
          }
        }
        if (!leaf) return false;
        const uint32_t t = link & ~LINK_LEAF;
        if (MIXED && near_triangle(a.w, l2, near_d)) {
          if (nn < (uint32_t)kRingDefer) near[nn * 64] = t;
          nn = nn < (uint32_t)kRingDefer ? nn + 1 : (uint32_t)kRingDefer + 1;
        } else {
          const float* p = tri + 9 * (size_t)t;
          p32 = __fadd_rn(p32, tri_solid_angle_fast(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], qx, qy,
                                                    qz));
        }
        return true;
      });
      const float s32 = wave_sum(p32);
      if (!MIXED) {
        w = (double)s32 / (4.0 * PI_D);
      } else if (__ballot(nn > (uint32_t)kRingDefer) != 0) {
        w = __longlong_as_double(0x7FF8000000000000ll);
      } else {
        double p64 = 0.0;
        for (uint32_t k = 0; k < nn; ++k) {
          const float* p = tri + 9 * (size_t)near[k * 64];
          p64 = __dadd_rn(p64, tri_solid_angle<double>(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], qx, qy,
                                                       qz));
        }
        w = ((double)s32 + wave_sum(p64)) / (4.0 * PI_D);
      }
    }
    if (lane == 0) W[qi] = w;
  }
}

// Mixed form after k_walk_ring: queries whose W falls in the ambiguous band
// are walked again in FP64, one wavefront per query (the others exit at once).
// A kernel of its own so that the FP64 walk's registers do not limit the
// occupancy of the FP32 walk.
template <int ORDER>
__global__ __launch_bounds__(256) void k_reeval64_band(const float4* __restrict__ slots,
                                                       const double* __restrict__ s64, uint32_t nslots,
                                                       const float* __restrict__ tri, const float* __restrict__ Q,
                                                       int64_t nq, double acc2, double band, double* __restrict__ W) {
  __shared__ uint2 ring_all[4][kReevalRing];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t qi = (int64_t)blockIdx.x * 4 + wv; qi < nq; qi += (int64_t)gridDim.x * 4) {
    if (fabs(W[qi] - 0.5) >= band) continue;  // wave-uniform; NaN (near-list overflow) is walked again
    const double w = ring_walk64<ORDER>(ring_all[wv], lane, slots, s64, nslots, tri, Q[3 * qi], Q[3 * qi + 1],
                                        Q[3 * qi + 2], acc2);
    if (lane == 0) W[qi] = w;
  }
}

// queries with |w - 0.5| < band, and queries whose walk gave NaN (more near
// triangles than the walk notes), are collected for the FP64 re-evaluation
__global__ void k_collect_ambiguous(const double* __restrict__ W, int64_t nq, double band, uint32_t* __restrict__ list,
                                    uint32_t* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const double w = W[i];
  if (isnan(w) || fabs(w - 0.5) < band) {
    const uint32_t p = atomicAdd(count, 1u);
    list[p] = (uint32_t)i;
  }
}

// Morton key of a query (30 bits) over the mesh bbox, for query ordering
__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
__global__ void k_morton_q(const float* __restrict__ Q, int64_t nq, float lx, float ly, float lz, float sx, float sy,
                           float sz, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  float fx = (Q[3 * i] - lx) * sx, fy = (Q[3 * i + 1] - ly) * sy, fz = (Q[3 * i + 2] - lz) * sz;
  fx = fminf(fmaxf(fx, 0.0f), 1023.0f);
  fy = fminf(fmaxf(fy, 0.0f), 1023.0f);
  fz = fminf(fmaxf(fz, 0.0f), 1023.0f);
  keys[i] = (expand_bits10((uint32_t)fx) << 2) | (expand_bits10((uint32_t)fy) << 1) | expand_bits10((uint32_t)fz);
  vals[i] = (uint32_t)i;
}

// 30-bit Hilbert key of a query (10 bits per axis, same quantisation as the
// Morton key), by Skilling's transpose algorithm ("Programming the Hilbert
// curve", AIP Conf. Proc. 707, 2004): undo the excess work, Gray-encode,
// interleave.  Consecutive keys are always neighbouring cells (the Morton
// curve jumps), so a wavefront of consecutive queries spans a more compact
// region: a CPU simulation of the wave walk on the 81,920-triangle workload
// at accuracy_scale 4.9 gave 1,471 record steps per wavefront against 1,544.
__device__ __forceinline__ uint32_t hilbert3_10(uint32_t x0, uint32_t x1, uint32_t x2) {
  uint32_t X[3] = {x0, x1, x2};
  // branch-free: m is all ones where bit q of X[i] is set
#pragma unroll
  for (int b = 9; b >= 1; --b) {
    const uint32_t p = (1u << b) - 1u;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const uint32_t m = 0u - ((X[i] >> b) & 1u);
      X[0] ^= p & m;
      const uint32_t t = (X[0] ^ X[i]) & p & ~m;
      X[0] ^= t;
      X[i] ^= t;
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
#pragma unroll
  for (int b = 9; b >= 1; --b) t ^= ((1u << b) - 1u) & (0u - ((X[2] >> b) & 1u));
  X[0] ^= t;
  X[1] ^= t;
  X[2] ^= t;
  uint32_t h = 0;
#pragma unroll
  for (int b = 9; b >= 0; --b) h = (h << 3) | (((X[0] >> b) & 1u) << 2) | (((X[1] >> b) & 1u) << 1) | ((X[2] >> b) & 1u);
  return h;
}
__global__ void k_hilbert_q(const float* __restrict__ Q, int64_t nq, float lx, float ly, float lz, float sx,
                            float sy, float sz, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  float fx = (Q[3 * i] - lx) * sx, fy = (Q[3 * i + 1] - ly) * sy, fz = (Q[3 * i + 2] - lz) * sz;
  fx = fminf(fmaxf(fx, 0.0f), 1023.0f);
  fy = fminf(fmaxf(fy, 0.0f), 1023.0f);
  fz = fminf(fmaxf(fz, 0.0f), 1023.0f);
  keys[i] = hilbert3_10((uint32_t)fx, (uint32_t)fy, (uint32_t)fz);
  vals[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// host orchestration
void compute_levels(const uint32_t* nodes, int64_t nn, Levels& L, hipStream_t st) {
  DevBuf<int32_t> parent(nn), depth(nn);
  DevBuf<uint32_t> hist(4097);
  DevBuf<int32_t> maxd(1);
  GWN_HIP(hipMemsetAsync(hist.p, 0, hist.bytes(), st));
  GWN_HIP(hipMemsetAsync(maxd.p, 0, sizeof(int32_t), st));
  const unsigned g = div_up(nn, 256);
  hipLaunchKernelGGL(k_parent, dim3(g), dim3(256), 0, st, nodes, nn, parent.p);
  check_launch("k_parent");
  hipLaunchKernelGGL(k_depth, dim3(g), dim3(256), 0, st, parent.p, nn, depth.p, hist.p, maxd.p);
  check_launch("k_depth");
  int32_t md = 0;
  GWN_HIP(hipMemcpyAsync(&md, maxd.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  GWN_CHECK(md < 4000, GWN_ERR_INVALID, "tree too deep");
  std::vector<uint32_t> h(md + 1);
  GWN_HIP(hipMemcpyAsync(h.data(), hist.p, sizeof(uint32_t) * (md + 1), hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  L.maxdepth = md;
  L.offset.assign(md + 2, 0);
  for (int d = 0; d <= md; ++d) L.offset[d + 1] = L.offset[d] + h[d];
  GWN_HIP(hipMemcpyAsync(hist.p, L.offset.data(), sizeof(uint32_t) * (md + 1), hipMemcpyHostToDevice, st));
  L.order.alloc(nn);
  hipLaunchKernelGGL(k_bucket, dim3(g), dim3(256), 0, st, depth.p, nn, hist.p, L.order.p);
  check_launch("k_bucket");
  GWN_HIP(hipStreamSynchronize(st));
}

// First slot of every node in the rope layout, top-down from the root's 0.
void rope_first_slots(const Levels& L, const uint32_t* nodes, const uint32_t* nsub, uint32_t* first,
                      hipStream_t st) {
  GWN_HIP(hipMemsetAsync(first, 0, sizeof(uint32_t), st));
  for (int d = 0; d <= L.maxdepth; ++d) {
    const int64_t cnt = L.offset[d + 1] - L.offset[d];
    if (!cnt) continue;
    hipLaunchKernelGGL(k_positions, dim3(div_up(cnt, 256)), dim3(256), 0, st, L.order.p + L.offset[d], cnt, nodes,
                       nsub, first);
    check_launch("k_positions");
  }
}

template <typename T>
static void upward(gwn_soup* s, const Levels& L, T* boxdata, DevBuf<uint32_t>& nsub, hipStream_t st) {
  const int64_t nn = s->nnodes;
  DevBuf<Local<T>> local(nn);
  for (int d = L.maxdepth; d >= 0; --d) {
    const int64_t cnt = L.offset[d + 1] - L.offset[d];
    if (!cnt) continue;
    hipLaunchKernelGGL(k_upward<T>, dim3(div_up(cnt, 128)), dim3(128), 0, st, L.order.p + L.offset[d], cnt,
                       s->nodes.p, s->U.p, s->F.p, s->boxes.p, s->order, local.p, boxdata, nsub.p);
    check_launch("k_upward");
  }
  GWN_HIP(hipStreamSynchronize(st));
}

// FP64 moments of every slot (the reference's item()/post() in double), in
// the field order of k_write_slots64: the all-FP64 form and the FP64
// re-evaluation of the mixed form read them.
static void build_fp64_moments(gwn_soup* s, const Levels& L, hipStream_t st) {
  const int64_t nn = s->nnodes;
  DevBuf<double> bd64(92 * nn);
  DevBuf<uint32_t> nsub(nn), first(nn);
  upward<double>(s, L, bd64.p, nsub, st);
  rope_first_slots(L, s->nodes.p, nsub.p, first.p, st);
  s->slots64.alloc(24 * (size_t)s->nslots);
  hipLaunchKernelGGL(k_write_slots64, dim3(div_up(nn, 256)), dim3(256), 0, st, nn, s->nodes.p, bd64.p, first.p,
                     nsub.p, s->slots64.p);
  check_launch("k_write_slots64");
  GWN_HIP(hipStreamSynchronize(st));
}

static void build_cost_grid(gwn_soup* s, hipStream_t st);

// Tree levels, FP32 BoxData (the reference's), the rope-ordered slot records,
// their FP64 moments and the cost grid.  After this the handle is read-only
// to queries.
void soup_finish_build(gwn_soup* s, hipStream_t st) {
  const int64_t nn = s->nnodes;
  if (nn == 0) {
    s->nslots = 0;
    return;
  }
  Levels L;
  compute_levels(s->nodes.p, nn, L, st);
  s->depth = L.maxdepth;
  s->boxdata.alloc(92 * nn);
  DevBuf<uint32_t> nsub(nn), first(nn);
  upward<float>(s, L, s->boxdata.p, nsub, st);
  uint32_t total = 0;
  GWN_HIP(hipMemcpyAsync(&total, nsub.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  rope_first_slots(L, s->nodes.p, nsub.p, first.p, st);
  GWN_HIP(hipStreamSynchronize(st));
  s->nslots = total;
  s->slots.alloc(6 * (size_t)total);
  hipLaunchKernelGGL(k_write_slots, dim3(div_up(nn, 256)), dim3(256), 0, st, nn, s->nodes.p, s->boxdata.p, nsub.p,
                     first.p, s->slots.p);
  check_launch("k_write_slots");
  build_fp64_moments(s, L, st);
  build_cost_grid(s, st);
  GWN_HIP(hipStreamSynchronize(st));
}

// Slot boxes for closest-point queries (and the antipodal rays), built by the
// first call that needs them.  Concurrent first calls on one handle build
// them once: the others wait on dist_mu, and the boxes are complete on the
// device (the build ends with a stream synchronisation) before dist_ready is
// set.
void soup_ensure_dist(gwn_soup* s, hipStream_t st) {
  if (s->dist_ready.load(std::memory_order_acquire) || s->nnodes == 0) return;
  std::lock_guard<std::mutex> lock(s->dist_mu);
  if (s->dist_ready.load(std::memory_order_relaxed)) return;
  const int64_t nn = s->nnodes;
  Levels L;
  compute_levels(s->nodes.p, nn, L, st);
  DevBuf<float> nbox(6 * (size_t)nn);
  for (int d = L.maxdepth; d >= 0; --d) {
    const int64_t cnt = L.offset[d + 1] - L.offset[d];
    if (!cnt) continue;
    hipLaunchKernelGGL(k_node_boxes, dim3(div_up(cnt, 128)), dim3(128), 0, st, L.order.p + L.offset[d], cnt,
                       s->nodes.p, s->V64.p, s->F.p, nbox.p);
    check_launch("k_node_boxes");
  }
  // first slot of every node, top-down, following the links of the rope
  DevBuf<uint32_t> first(nn);
  GWN_HIP(hipMemsetAsync(first.p, 0, sizeof(uint32_t), st));
  for (int d = 0; d <= L.maxdepth; ++d) {
    const int64_t cnt = L.offset[d + 1] - L.offset[d];
    if (!cnt) continue;
    hipLaunchKernelGGL(k_positions_rope, dim3(div_up(cnt, 256)), dim3(256), 0, st, L.order.p + L.offset[d], cnt,
                       s->nodes.p, s->slots.p, first.p);
    check_launch("k_positions_rope");
  }
  DevBuf<float4> db(2 * (size_t)s->nslots);
  hipLaunchKernelGGL(k_write_dbox, dim3(div_up(nn, 256)), dim3(256), 0, st, nn, s->nodes.p, nbox.p, s->V64.p,
                     s->F.p, first.p, s->slots.p, db.p);
  check_launch("k_write_dbox");
  GWN_HIP(hipStreamSynchronize(st));
  s->dbox = std::move(db);
  s->dist_ready.store(true, std::memory_order_release);
}

void soup_setup(gwn_soup* s, hipStream_t st) {
  const int64_t nf = s->nf;
  s->boxes.alloc(6 * (size_t)nf);
  s->tri.alloc(9 * (size_t)nf);
  if (nf > 0) {
    hipLaunchKernelGGL(k_tri_setup, dim3(div_up(nf, 256)), dim3(256), 0, st, s->U.p, s->F.p, nf, s->boxes.p,
                       s->tri.p);
    check_launch("k_tri_setup");
  }
}

// Workspace of one sorted query call: Hilbert keys, the query permutation,
// sort temp, and the list and count of ambiguous queries (mixed form).
size_t soup_query_work_bytes(int64_t nq) {
  const size_t a = 256;
  auto up = [&](size_t b) { return (b + a - 1) / a * a; };
  return up(4 * (size_t)nq) * 2 + up(radix_sort_temp_bytes(nq, false)) + up(4 * (size_t)nq) + up(4);
}
// Workspace of an unsorted (per-lane) call: the ambiguous list and count only.

// 30-bit Morton keys of the queries over a box (queries outside are clamped
// to it), and the identity permutation to be sorted along with them
void morton_keys(const float* Q, int64_t nq, const float* bb, uint32_t* keys, uint32_t* vals, hipStream_t st) {
  const float ex = std::max(bb[3] - bb[0], 1e-30f), ey = std::max(bb[4] - bb[1], 1e-30f),
              ez = std::max(bb[5] - bb[2], 1e-30f);
  hipLaunchKernelGGL(k_morton_q, dim3(div_up(nq, 256)), dim3(256), 0, st, Q, nq, bb[0], bb[1], bb[2], 1024.0f / ex,
                     1024.0f / ey, 1024.0f / ez, keys, vals);
  check_launch("k_morton_q");
}


// ---------------------------------------------------------------------------
// Block order of the wave walks.  A wavefront's walk time is the union of its
// 64 queries' paths, and it varies several-fold over a batch (queries near the
// surface open many more records); dispatched in Hilbert order, the costly
// blocks of a region can start last and run on alone at the end of the
// launch.  The walks instead take each XCD's blocks costliest first (longest
// processing time first), with a block's cost estimated from the cost grid
// built with the tree.  The order changes which blocks run when, never which
// queries a wavefront holds, so every result is the same bit for bit.

// Records a query at each cell centre of the cost grid visits at accuracy
// scale 2 (the near/far decisions of k_traverse, without the sums), as a
// log-scale byte: min(255, 15 log2(1 + n)).
__global__ __launch_bounds__(256) void k_cost_grid(const float4* __restrict__ slots, uint32_t nslots, float lx,
                                                   float ly, float lz, float hx, float hy, float hz,
                                                   uint8_t* __restrict__ grid) {
  constexpr int G = kCostGrid;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G * G * G) return;
  const float qx = lx + (c % G + 0.5f) * hx, qy = ly + ((c / G) % G + 0.5f) * hy, qz = lz + (c / (G * G) + 0.5f) * hz;
  uint32_t i = 0, n = 0;
  while (i < nslots) {
    const float4 a = slots[6 * (size_t)i];
    const uint32_t link = __float_as_uint(slots[6 * (size_t)i + 1].w);
    const bool leaf = (link & LINK_LEAF) != 0;
    const float x = qx - a.x, y = qy - a.y, z = qz - a.z;
    const float l2 = x * x + y * y + z * z;
    ++n;
    if (!(l2 <= a.w * 4.0f)) {
      i = leaf ? i + 1 : link;
    } else {
      n += leaf;
      ++i;
    }
  }
  grid[c] = (uint8_t)fminf(255.0f, 15.0f * log2f(1.0f + (float)n));
}

static void build_cost_grid(gwn_soup* s, hipStream_t st) {
  constexpr int G = kCostGrid;
  s->cost_grid.alloc((size_t)G * G * G);
  const float* bb = s->bbox;
  hipLaunchKernelGGL(k_cost_grid, dim3(div_up(G * G * G, 256)), dim3(256), 0, st, s->slots.p, (uint32_t)s->nslots,
                     bb[0], bb[1], bb[2], (bb[3] - bb[0]) / G, (bb[4] - bb[1]) / G, (bb[5] - bb[2]) / G,
                     s->cost_grid.p);
  check_launch("k_cost_grid");
}

// Estimated cost of each block of 256 sorted queries: the grid's largest
// value over the block's queries.
__global__ __launch_bounds__(256) void k_block_cost(const float* __restrict__ Q, const uint32_t* __restrict__ perm,
                                                    int64_t nq, float lx, float ly, float lz, float sx, float sy,
                                                    float sz, const uint8_t* __restrict__ grid,
                                                    uint8_t* __restrict__ cost) {
  constexpr int G = kCostGrid;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int v = 0;
  if (t < nq) {
    const int64_t qi = perm[t];
    const int ix = min(G - 1, max(0, (int)((Q[3 * qi] - lx) * sx)));
    const int iy = min(G - 1, max(0, (int)((Q[3 * qi + 1] - ly) * sy)));
    const int iz = min(G - 1, max(0, (int)((Q[3 * qi + 2] - lz) * sz)));
    v = grid[(iz * G + iy) * G + ix];
  }
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
  __shared__ int wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) cost[blockIdx.x] = (uint8_t)max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
}

// Block order: one workgroup per XCD's range of logical blocks (the ranges of
// xcd_block()), a counting sort of its blocks by decreasing cost.  Blocks of
// equal cost land in any order.
__global__ __launch_bounds__(256) void k_block_order(const uint8_t* __restrict__ cost, uint32_t nb,
                                                     uint32_t* __restrict__ order) {
  __shared__ uint32_t count[256];
  const uint32_t x = blockIdx.x, q = nb / 8, r = nb % 8;
  const uint32_t lo = x * q + (x < r ? x : r), hi = lo + q + (x < r ? 1 : 0);
  count[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t b = lo + threadIdx.x; b < hi; b += blockDim.x) atomicAdd(&count[255 - cost[b]], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = lo;
    for (int k = 0; k < 256; ++k) {
      const uint32_t n = count[k];
      count[k] = run;
      run += n;
    }
  }
  __syncthreads();
  for (uint32_t b = lo + threadIdx.x; b < hi; b += blockDim.x) order[atomicAdd(&count[255 - cost[b]], 1u)] = b;
}

// order (nb = ceil(nq / 256) entries) and cost (nb bytes) are workspace
void block_order(const uint8_t* grid, const float* bb, const float* Q, const uint32_t* perm, int64_t nq,
                 uint32_t* order, uint8_t* cost, hipStream_t st) {
  constexpr int G = kCostGrid;
  const uint32_t nb = (uint32_t)div_up(nq, 256);
  const float ex = std::max(bb[3] - bb[0], 1e-30f), ey = std::max(bb[4] - bb[1], 1e-30f),
              ez = std::max(bb[5] - bb[2], 1e-30f);
  hipLaunchKernelGGL(k_block_cost, dim3(nb), dim3(256), 0, st, Q, perm, nq, bb[0], bb[1], bb[2], G / ex, G / ey,
                     G / ez, grid, cost);
  check_launch("k_block_cost");
  hipLaunchKernelGGL(k_block_order, dim3(8), dim3(256), 0, st, cost, nb, order);
  check_launch("k_block_order");
}

// 30-bit Hilbert keys of the queries over a box (clamped to it) and the
// identity permutation: the order in which the tree walks take their queries
void hilbert_keys(const float* Q, int64_t nq, const float* bb, uint32_t* keys, uint32_t* vals, hipStream_t st) {
  const float ex = std::max(bb[3] - bb[0], 1e-30f), ey = std::max(bb[4] - bb[1], 1e-30f),
              ez = std::max(bb[5] - bb[2], 1e-30f);
  hipLaunchKernelGGL(k_hilbert_q, dim3(div_up(nq, 256)), dim3(256), 0, st, Q, nq, bb[0], bb[1], bb[2], 1024.0f / ex,
                     1024.0f / ey, 1024.0f / ez, keys, vals);
  check_launch("k_hilbert_q");
}

// ---------------------------------------------------------------------------
// Scratch-free query (gwn_soup_query_device_noalloc).  Its workspace is this
// file-scope device array, not pooled memory: each device that runs the
// library holds one copy for the life of the process.  Host code never uses
// the array's name as an address; it uses the device address that
// hipGetSymbolAddress returns, looked up once per device (with that device
// current) on the first call there.  Every call on a device shares that
// device's copy, so calls that overlap on one device race; the contents a
// call needs zeroed (the ambiguous-query count) are zeroed by a kernel on the
// call's stream.
constexpr int64_t kNoAllocMaxQueries = (int64_t)1 << 20;
constexpr size_t kNoAllocScratchBytes = (size_t)22 << 20;  // >= soup_query_work_bytes(kNoAllocMaxQueries)
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

static uint8_t* noalloc_scratch() {
  static std::mutex mu;
  static uint8_t* ptr[kNoAllocMaxDevices] = {};
  int dev = 0;
  GWN_HIP(hipGetDevice(&dev));
  GWN_CHECK(dev >= 0 && dev < kNoAllocMaxDevices, GWN_ERR_NODEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      throw Error{GWN_ERR_HIP, std::string("hipGetSymbolAddress of the query scratch failed: ") + hipGetErrorString(e)};
    }
    GWN_CHECK(soup_query_work_bytes(kNoAllocMaxQueries) <= kNoAllocScratchBytes, GWN_ERR_INVALID,
              "scratch-free query: workspace layout exceeds the static scratch");
    ptr[dev] = static_cast<uint8_t*>(p);
  }
  return ptr[dev];
}

int64_t soup_noalloc_max_queries() { return kNoAllocMaxQueries; }

int soup_walk_for(const gwn_soup* s, int64_t nq, bool latency_bound) {
  if (nq < kRingMaxQueries) return WALK_RING;
  return latency_bound && s->nslots >= (1 << 17) ? WALK_WAVE_HANDOVER : WALK_WAVE;
}

void soup_query_device(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                       double* W, hipStream_t st) {
  if (nq <= 0) return;
  const int walk = soup_walk_for(s, nq, false);
  if (walk == WALK_RING) {
    soup_query_device_ws(s, Q, nq, accuracy_scale, precision, W, st, nullptr, 0, WALK_RING);
    return;
  }
  Scratch ws(soup_query_work_bytes(nq), st);
  soup_query_device_ws(s, Q, nq, accuracy_scale, precision, W, st, ws.p, ws.capacity, walk);
}

// The scratch-free form of soup_query_device: same kernels, same results;
// the workspace is the static scratch, the mixed form's FP64 re-evaluation is
// sized on the device (no host round trip) and no timing events are made.
void soup_query_device_noalloc(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                               double* W, hipStream_t st) {
  GWN_CHECK(nq <= kNoAllocMaxQueries, GWN_ERR_INVALID,
            "scratch-free query: at most 1,048,576 queries per call (got " + std::to_string(nq) + ")");
  if (nq <= 0) return;
  const int walk = soup_walk_for(s, nq, false);
  if (walk == WALK_RING) {
    soup_query_device_ws(s, Q, nq, accuracy_scale, precision, W, st, nullptr, 0, WALK_RING, WS_NO_PROFILE);
    return;
  }
  uint8_t* ws = noalloc_scratch();
  soup_query_device_ws(s, Q, nq, accuracy_scale, precision, W, st, ws, kNoAllocScratchBytes, walk,
                       WS_DEVICE_COUNT | WS_NO_PROFILE);
}

// FP64 walk over nq queries in the order perm
static void launch_traverse64(const gwn_soup* s, unsigned g, hipStream_t st, const float* Q, const uint32_t* perm,
                              const uint32_t* order, int64_t nq, double a2, double* W) {
  const uint32_t ns = (uint32_t)s->nslots;
  switch (s->order) {
    case 0: hipLaunchKernelGGL(k_traverse64<0>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, perm, order, nq, nullptr, a2, W); break;
    case 1: hipLaunchKernelGGL(k_traverse64<1>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, perm, order, nq, nullptr, a2, W); break;
    default: hipLaunchKernelGGL(k_traverse64<2>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, perm, order, nq, nullptr, a2, W); break;
  }
  check_launch("k_traverse64");
}

// one wavefront per query (k_walk_ring), 4 per block
static void launch_walk_ring(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                             double* W, hipStream_t st, bool prof_on) {
  const uint32_t ns = (uint32_t)s->nslots;
  const float acc2 = accuracy_scale * accuracy_scale;
  const double a2 = (double)accuracy_scale * (double)accuracy_scale;
  const float nd = s->mixed.near_dist;
  const double band = (double)s->mixed.ambiguous_band;
  const unsigned g = (unsigned)div_up(nq, 4);
  ProfileScope prof("soup_walk_ring", st, prof_on);
  if (precision == GWN_PREC_FP64) {
    switch (s->order) {
      case 0: hipLaunchKernelGGL(k_reeval64<0>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, nullptr, nq, nullptr, a2, W); break;
      case 1: hipLaunchKernelGGL(k_reeval64<1>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, nullptr, nq, nullptr, a2, W); break;
      default: hipLaunchKernelGGL(k_reeval64<2>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, nullptr, nq, nullptr, a2, W); break;
    }
    check_launch("k_reeval64");
    return;
  }
#define GWN_LAUNCH_R(O, M)                                                                                      \
  hipLaunchKernelGGL((k_walk_ring<O, M>), dim3(g), dim3(256), 0, st, s->slots.p, ns, s->tri.p, Q, nq, acc2, nd, W)
  const bool mixed = precision == GWN_PREC_MIXED;
  switch (s->order) {
    case 0: if (mixed) GWN_LAUNCH_R(0, true); else GWN_LAUNCH_R(0, false); break;
    case 1: if (mixed) GWN_LAUNCH_R(1, true); else GWN_LAUNCH_R(1, false); break;
    default: if (mixed) GWN_LAUNCH_R(2, true); else GWN_LAUNCH_R(2, false); break;
  }
#undef GWN_LAUNCH_R
  check_launch("k_walk_ring");
  if (mixed) {
    switch (s->order) {
      case 0: hipLaunchKernelGGL(k_reeval64_band<0>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, nq, a2, band, W); break;
      case 1: hipLaunchKernelGGL(k_reeval64_band<1>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, nq, a2, band, W); break;
      default: hipLaunchKernelGGL(k_reeval64_band<2>, dim3(g), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, nq, a2, band, W); break;
    }
    check_launch("k_reeval64_band");
  }
}

// Query with caller-provided device workspace of soup_query_work_bytes(nq)
// bytes (none for WALK_RING); calls with distinct workspaces may run
// concurrently on distinct streams against the same tree.  flags:
// WS_DEVICE_COUNT launches the mixed form's FP64 re-evaluation over every
// query with the count read on the device (no host round trip); WS_NO_PROFILE
// creates no kernel-timing events.
void soup_query_device_ws(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                          double* W, hipStream_t st, uint8_t* workspace, size_t ws_bytes, int walk, int flags) {
  if (nq <= 0) return;
  const bool prof_on = (flags & WS_NO_PROFILE) == 0;
  if (s->nslots == 0) {
    fill_device<double>(W, nq, 0.0, st);
    return;
  }
  if (walk == WALK_RING) {
    launch_walk_ring(s, Q, nq, accuracy_scale, precision, W, st, prof_on);
    return;
  }
  GWN_CHECK(walk == WALK_WAVE || walk == WALK_WAVE_HANDOVER, GWN_ERR_INVALID, "unknown walk");
  GWN_CHECK(ws_bytes >= soup_query_work_bytes(nq), GWN_ERR_INVALID, "query workspace too small");
  uint8_t* w = workspace;
  auto take = [&](size_t b) {
    uint8_t* p = w;
    w += (b + 255) / 256 * 256;
    return p;
  };
  uint32_t* keys = (uint32_t*)take(4 * (size_t)nq);
  uint32_t* perm = (uint32_t*)take(4 * (size_t)nq);
  {
    const size_t tmpb = radix_sort_temp_bytes(nq, false);
    void* tmp = take(tmpb);
    hilbert_keys(Q, nq, s->bbox, keys, perm, st);
    ProfileScope prof("radix_sort", st, prof_on);
    radix_sort_u32(keys, perm, nq, 30, tmp, tmpb, st);
  }
  uint32_t* amb = (uint32_t*)take(4 * (size_t)nq);
  uint32_t* ambn = (uint32_t*)take(4);
  const unsigned g = div_up(nq, 256);
  // the sort keys are spent: their space holds the block order (g entries)
  // and the block costs (g bytes), 5g <= 4 nq bytes rounded up to 256
  uint32_t* border = keys;
  {
    ProfileScope prof("block_order", st, prof_on);
    block_order(s->cost_grid.p, s->bbox, Q, perm, nq, border, (uint8_t*)(border + g), st);
  }

  const float acc2 = accuracy_scale * accuracy_scale;
  const uint32_t ns = (uint32_t)s->nslots;
  if (precision == GWN_PREC_FP64) {
    const double a2 = (double)accuracy_scale * (double)accuracy_scale;
    ProfileScope prof("soup_traverse64", st, prof_on);
    launch_traverse64(s, g, st, Q, perm, border, nq, a2, W);
    return;
  }
  const bool mixed = precision == GWN_PREC_MIXED;
  const float nd = s->mixed.near_dist;
#define GWN_LAUNCH_T(O, M)                                                                                      \
  do {                                                                                                          \
    if (walk == WALK_WAVE)                                                                                      \
      hipLaunchKernelGGL((k_traverse<O, M, false>), dim3(g), dim3(256), 0, st, s->slots.p, ns, s->tri.p, Q, perm, \
                         border, nq, acc2, nd, W);                                                              \
    else                                                                                                        \
      hipLaunchKernelGGL((k_traverse<O, M, true>), dim3(g), dim3(256), 0, st, s->slots.p, ns, s->tri.p, Q, perm, \
                         border, nq, acc2, nd, W);                                                              \
  } while (0)
  {
    ProfileScope prof("soup_traverse", st, prof_on);
    if (mixed) {
      switch (s->order) {
        case 0: GWN_LAUNCH_T(0, true); break;
        case 1: GWN_LAUNCH_T(1, true); break;
        default: GWN_LAUNCH_T(2, true); break;
      }
    } else {
      switch (s->order) {
        case 0: GWN_LAUNCH_T(0, false); break;
        case 1: GWN_LAUNCH_T(1, false); break;
        default: GWN_LAUNCH_T(2, false); break;
      }
    }
  }
#undef GWN_LAUNCH_T
  check_launch("k_traverse");
  if (mixed) {
    fill_device<uint32_t>(ambn, 1, 0u, st);
    hipLaunchKernelGGL(k_collect_ambiguous, dim3(g), dim3(256), 0, st, W, nq, (double)s->mixed.ambiguous_band, amb,
                       ambn);
    check_launch("k_collect_ambiguous");
    const double a2 = (double)accuracy_scale * (double)accuracy_scale;
    int64_t na64 = 0;
    const uint32_t* dcount = nullptr;
    if (flags & WS_DEVICE_COUNT) {
      // bounded by every query; the kernel reads the count on the device
      na64 = nq;
      dcount = ambn;
    } else {
      // the count comes back to the host so that the FP64 pass is launched at
      // its size, or not at all
      uint32_t na = 0;
      GWN_HIP(hipMemcpyAsync(&na, ambn, 4, hipMemcpyDeviceToHost, st));
      GWN_HIP(hipStreamSynchronize(st));
      na64 = na;
    }
    if (na64 > 0) {
      // one wavefront per query, four per block; with the count on the
      // device the grid is capped and strides over the list
      const unsigned gr = dcount ? (unsigned)std::min<int64_t>(div_up(na64, 4), 2048) : (unsigned)div_up(na64, 4);
      ProfileScope prof("soup_reeval64", st, prof_on);
      switch (s->order) {
        case 0: hipLaunchKernelGGL(k_reeval64<0>, dim3(gr), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, amb, na64, dcount, a2, W); break;
        case 1: hipLaunchKernelGGL(k_reeval64<1>, dim3(gr), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, amb, na64, dcount, a2, W); break;
        default: hipLaunchKernelGGL(k_reeval64<2>, dim3(gr), dim3(256), 0, st, s->slots.p, s->slots64.p, ns, s->tri.p, Q, amb, na64, dcount, a2, W); break;
      }
      check_launch("k_reeval64");
    }
  }
}

}  // namespace gwn
