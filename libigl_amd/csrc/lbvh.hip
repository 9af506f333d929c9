// lbvh.hip -- 4-wide BVH built on the GPU from Morton codes.
//
//   k_lbvh_keys    triangles whose box has a NaN/inf coordinate are dropped
//                  (as UT::BVH::init drops them, FastWindingNumberForSoups.h:
//                  4535-4615); the others get a 30-bit Morton code of their box
//                  centre in the scene box (10 bits per axis)
//   radix sort     stable LDS radix sort of (code, triangle) -- ties keep the
//                  triangle order, so the permutation equals a stable CPU sort
//   k_karras       binary radix tree over the sorted codes (Karras 2012): one
//                  thread per internal node finds its range by doubling and
//                  binary search on common-prefix lengths; equal codes are
//                  separated by the position (delta = 32 + clz(i ^ j))
//   k_bin_depth    depth of every internal node (walk to the root)
//   k_collapse     every internal node at even depth becomes a 4-wide node
//                  whose children are its grandchildren (or its children where
//                  those are leaves), in left-to-right Morton order; nodes are
//                  numbered by an exclusive scan of the even-depth flags, so
//                  the root (binary node 0) is node 0.
// The output uses the reference's node encoding (item index, 0x80000000|node,
// 0xFFFFFFFF empty), so the moment pass and traversal are shared with the
// reference-topology tree.
#include <hip/hip_runtime.h>

#include "gwn_internal.h"

namespace gwn {

__device__ __forceinline__ uint32_t lb_expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ bool lb_bad(float x) { return (__float_as_uint(x) & 0x7F800000u) == 0x7F800000u; }

__global__ void k_lbvh_flags(const float* __restrict__ boxes, int64_t nf, uint32_t* __restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  bool bad = false;
  for (int k = 0; k < 6; ++k) bad |= lb_bad(boxes[6 * i + k]);
  keep[i] = bad ? 0u : 1u;
}

// Morton code of the box centre; the same float steps as the CPU restatement
// (centre = 0.5*(lo+hi), t = (centre - lo_scene) * scale, floor, clamp 0..1023).
__global__ void k_lbvh_keys(const float* __restrict__ boxes, int64_t nf, const uint32_t* __restrict__ keep,
                            const uint32_t* __restrict__ slot, float lx, float ly, float lz, float sx, float sy,
                            float sz, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf || !keep[i]) return;
  const float* b = boxes + 6 * i;
  const float cx = 0.5f * (b[0] + b[1]), cy = 0.5f * (b[2] + b[3]), cz = 0.5f * (b[4] + b[5]);
  float tx = (cx - lx) * sx, ty = (cy - ly) * sy, tz = (cz - lz) * sz;
  tx = fminf(fmaxf(tx, 0.0f), 1023.0f);
  ty = fminf(fmaxf(ty, 0.0f), 1023.0f);
  tz = fminf(fmaxf(tz, 0.0f), 1023.0f);
  const uint32_t code =
      (lb_expand10((uint32_t)tx) << 2) | (lb_expand10((uint32_t)ty) << 1) | lb_expand10((uint32_t)tz);
  const uint32_t p = slot[i];
  keys[p] = code;
  vals[p] = (uint32_t)i;
}

__device__ __forceinline__ int lb_delta(const uint32_t* __restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Binary node ids: internal i in [0, n-1) -> i; leaf k in [0, n) -> LEAF_TAG | k.
constexpr uint32_t LB_LEAF = 0x80000000u;

__global__ void k_karras(const uint32_t* __restrict__ keys, int n, uint32_t* __restrict__ left,
                         uint32_t* __restrict__ right, int32_t* __restrict__ parent) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (lb_delta(keys, n, i, i + 1) - lb_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = lb_delta(keys, n, i, i - d);
  int lmax = 2;
  while (lb_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (lb_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = lb_delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) / 2;
    if (lb_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int g = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const uint32_t L = (lo == g) ? (LB_LEAF | (uint32_t)g) : (uint32_t)g;
  const uint32_t R = (hi == g + 1) ? (LB_LEAF | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
  left[i] = L;
  right[i] = R;
  if (!(L & LB_LEAF)) parent[L] = i;
  if (!(R & LB_LEAF)) parent[R] = i;
  if (i == 0) parent[0] = -1;
}

__global__ void k_bin_depth(const int32_t* __restrict__ parent, int m, uint32_t* __restrict__ even) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int d = 0;
  int32_t p = parent[i];
  while (p >= 0 && d < 1 << 20) {
    ++d;
    p = parent[p];
  }
  even[i] = (d & 1) ? 0u : 1u;
}

__global__ void k_collapse(const uint32_t* __restrict__ left, const uint32_t* __restrict__ right, int m,
                           const uint32_t* __restrict__ even, const uint32_t* __restrict__ number,
                           const uint32_t* __restrict__ tri_of_leaf, uint32_t* __restrict__ nodes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !even[i]) return;
  uint32_t out[4] = {NODE_EMPTY, NODE_EMPTY, NODE_EMPTY, NODE_EMPTY};
  int c = 0;
  const uint32_t kids[2] = {left[i], right[i]};
  for (int a = 0; a < 2; ++a) {
    const uint32_t k = kids[a];
    if (k & LB_LEAF) {
      out[c++] = tri_of_leaf[k & ~LB_LEAF];
    } else {
      const uint32_t gk[2] = {left[k], right[k]};
      for (int b = 0; b < 2; ++b) {
        const uint32_t g = gk[b];
        out[c++] = (g & LB_LEAF) ? tri_of_leaf[g & ~LB_LEAF] : (NODE_INTERNAL | number[g]);
      }
    }
  }
  uint32_t* o = nodes + 4 * (size_t)number[i];
  for (int s = 0; s < 4; ++s) o[s] = out[s];
}

static size_t a256(size_t b) { return (b + 255) / 256 * 256; }

int64_t lbvh_build(const float* boxes, int64_t nf, const float* bbox, DevBuf<uint32_t>& nodes, hipStream_t st) {
  DevBuf<uint32_t> keep(nf), slot(nf), total(1);
  DevBuf<uint8_t> sc(scan_temp_bytes(nf));
  hipLaunchKernelGGL(k_lbvh_flags, dim3(div_up(nf, 256)), dim3(256), 0, st, boxes, nf, keep.p);
  check_launch("k_lbvh_flags");
  exclusive_scan_u32(keep.p, slot.p, nf, total.p, sc.p, st);
  uint32_t n = 0;
  GWN_HIP(hipMemcpyAsync(&n, total.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  if (n == 0) return 0;
  DevBuf<uint32_t> keys(n), vals(n);
  const float ex = bbox[3] - bbox[0], ey = bbox[4] - bbox[1], ez = bbox[5] - bbox[2];
  const float sx = ex > 0 ? 1024.0f / ex : 0.0f, sy = ey > 0 ? 1024.0f / ey : 0.0f, sz = ez > 0 ? 1024.0f / ez : 0.0f;
  hipLaunchKernelGGL(k_lbvh_keys, dim3(div_up(nf, 256)), dim3(256), 0, st, boxes, nf, keep.p, slot.p, bbox[0],
                     bbox[1], bbox[2], sx, sy, sz, keys.p, vals.p);
  check_launch("k_lbvh_keys");
  {
    const size_t tb = radix_sort_temp_bytes(n, false);
    DevBuf<uint8_t> tmp(tb);
    radix_sort_u32(keys.p, vals.p, n, 30, tmp.p, tb, st);
  }
  if (n == 1) {
    const uint32_t root[4] = {0, NODE_EMPTY, NODE_EMPTY, NODE_EMPTY};
    uint32_t tri = 0;
    GWN_HIP(hipMemcpyAsync(&tri, vals.p, 4, hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
    uint32_t r[4] = {tri, root[1], root[2], root[3]};
    nodes.alloc(4);
    GWN_HIP(hipMemcpyAsync(nodes.p, r, 16, hipMemcpyHostToDevice, st));
    GWN_HIP(hipStreamSynchronize(st));
    return 1;
  }
  const int m = (int)n - 1;
  DevBuf<uint32_t> L(m), R(m), even(m), number(m), cnt(1);
  DevBuf<int32_t> parent(m);
  hipLaunchKernelGGL(k_karras, dim3(div_up(m, 256)), dim3(256), 0, st, keys.p, (int)n, L.p, R.p, parent.p);
  check_launch("k_karras");
  hipLaunchKernelGGL(k_bin_depth, dim3(div_up(m, 256)), dim3(256), 0, st, parent.p, m, even.p);
  check_launch("k_bin_depth");
  DevBuf<uint8_t> sc2(scan_temp_bytes(m));
  exclusive_scan_u32(even.p, number.p, m, cnt.p, sc2.p, st);
  uint32_t nn = 0;
  GWN_HIP(hipMemcpyAsync(&nn, cnt.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  nodes.alloc(4 * (size_t)nn);
  hipLaunchKernelGGL(k_collapse, dim3(div_up(m, 256)), dim3(256), 0, st, L.p, R.p, m, even.p, number.p, vals.p,
                     nodes.p);
  check_launch("k_collapse");
  GWN_HIP(hipStreamSynchronize(st));
  (void)a256;
  return (int64_t)nn;
}

}  // namespace gwn
