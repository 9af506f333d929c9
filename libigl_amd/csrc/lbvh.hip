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
//   k_bin_boxes    box of every binary node, bottom-up: one thread per leaf
//                  walks to the root and the second arrival at a node joins
//                  its children's boxes (min/max, so the order of arrival
//                  does not change them)
//   k_expand /     the 4-wide nodes, top-down one level at a time: a node
//   k_emit         starts from its binary node's two children and opens its
//                  largest member (half surface area dx dy + dy dz + dz dx,
//                  first on ties) until it has four members or only leaves;
//                  members keep their left-to-right Morton order, and nodes
//                  are numbered level by level (the root is node 0).  On the
//                  81,920-triangle headline mesh this greedy collapse walked
//                  faster than taking every even-depth node's grandchildren.
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
                         uint32_t* __restrict__ right, int32_t* __restrict__ parent,
                         int32_t* __restrict__ leaf_parent) {
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
  if (L & LB_LEAF) leaf_parent[L & ~LB_LEAF] = i;
  else parent[L] = i;
  if (R & LB_LEAF) leaf_parent[R & ~LB_LEAF] = i;
  else parent[R] = i;
  if (i == 0) parent[0] = -1;
}

__device__ __forceinline__ void lb_join(const float* __restrict__ a, const float* __restrict__ b, float* o) {
  for (int k = 0; k < 3; ++k) {
    o[2 * k] = fminf(a[2 * k], b[2 * k]);
    o[2 * k + 1] = fmaxf(a[2 * k + 1], b[2 * k + 1]);
  }
}

// Boxes of the binary nodes (6 floats: lo/hi per axis, as the item boxes),
// bottom-up from the leaves; arrivals[] starts at zero.
__global__ void k_bin_boxes(const float* __restrict__ boxes, const uint32_t* __restrict__ tri_of_leaf, int n,
                            const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                            const int32_t* __restrict__ parent, const int32_t* __restrict__ leaf_parent,
                            uint32_t* __restrict__ arrivals, float* __restrict__ nbox) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t p = leaf_parent[k];
  while (p >= 0) {
    // this subtree's box is written; the sibling's is done when the other
    // thread arrived first
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (atomicAdd(&arrivals[p], 1u) == 0u) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t c[2] = {left[p], right[p]};
    float b[2][6];
    for (int s = 0; s < 2; ++s) {
      const float* src = (c[s] & LB_LEAF) ? boxes + 6 * (size_t)tri_of_leaf[c[s] & ~LB_LEAF] : nbox + 6 * (size_t)c[s];
      for (int t = 0; t < 6; ++t) b[s][t] = src[t];
    }
    lb_join(b[0], b[1], nbox + 6 * (size_t)p);
    p = parent[p];
  }
}

__device__ __forceinline__ float lb_area(const float* b) {
  const float dx = b[1] - b[0], dy = b[3] - b[2], dz = b[5] - b[4];
  return (dx * dy + dy * dz) + dz * dx;
}

// Members of each frontier node (binary node fr[j]): its two children, then
// the largest internal member opened in place until four or all leaves.
__global__ void k_expand(const uint32_t* __restrict__ fr, int F, const uint32_t* __restrict__ left,
                         const uint32_t* __restrict__ right, const float* __restrict__ nbox,
                         uint32_t* __restrict__ members, uint32_t* __restrict__ nint) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= F) return;
  const uint32_t b = fr[j];
  uint32_t mem[4] = {left[b], right[b], NODE_EMPTY, NODE_EMPTY};
  int c = 2;
  while (c < 4) {
    int best = -1;
    float ba = -1.0f;
    for (int t = 0; t < c; ++t)
      if (!(mem[t] & LB_LEAF)) {
        const float a = lb_area(nbox + 6 * (size_t)mem[t]);
        if (a > ba) {
          ba = a;
          best = t;
        }
      }
    if (best < 0) break;
    const uint32_t x = mem[best];
    for (int t = c; t > best + 1; --t) mem[t] = mem[t - 1];
    mem[best] = left[x];
    mem[best + 1] = right[x];
    ++c;
  }
  uint32_t k = 0;
  for (int t = 0; t < 4; ++t) {
    members[4 * (size_t)j + t] = mem[t];
    if (t < c && !(mem[t] & LB_LEAF)) ++k;
  }
  nint[j] = k;
}

// 4-wide node base + j; its internal members become the next level's
// frontier at pos[j].., numbered base_next + pos[j] + ...
__global__ void k_emit(int F, const uint32_t* __restrict__ members, const uint32_t* __restrict__ pos,
                       uint32_t base, uint32_t base_next, const uint32_t* __restrict__ tri_of_leaf,
                       uint32_t* __restrict__ nodes, uint32_t* __restrict__ next_fr) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= F) return;
  uint32_t k = pos[j];
  uint32_t* o = nodes + 4 * (size_t)(base + j);
  for (int t = 0; t < 4; ++t) {
    const uint32_t m = members[4 * (size_t)j + t];
    if (m == NODE_EMPTY) {
      o[t] = NODE_EMPTY;
    } else if (m & LB_LEAF) {
      o[t] = tri_of_leaf[m & ~LB_LEAF];
    } else {
      o[t] = NODE_INTERNAL | (base_next + k);
      next_fr[k++] = m;
    }
  }
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
  DevBuf<uint32_t> L(m), R(m), arrivals(m);
  DevBuf<int32_t> parent(m), leaf_parent(n);
  hipLaunchKernelGGL(k_karras, dim3(div_up(m, 256)), dim3(256), 0, st, keys.p, (int)n, L.p, R.p, parent.p,
                     leaf_parent.p);
  check_launch("k_karras");
  DevBuf<float> nbox(6 * (size_t)m);
  fill_device<uint32_t>(arrivals.p, m, 0u, st);
  hipLaunchKernelGGL(k_bin_boxes, dim3(div_up(n, 256)), dim3(256), 0, st, boxes, vals.p, (int)n, L.p, R.p, parent.p,
                     leaf_parent.p, arrivals.p, nbox.p);
  check_launch("k_bin_boxes");
  // top-down, one level of 4-wide nodes at a time (at most m nodes)
  DevBuf<uint32_t> tmp_nodes(4 * (size_t)m), fr[2] = {DevBuf<uint32_t>(m), DevBuf<uint32_t>(m)};
  DevBuf<uint32_t> members(4 * (size_t)m), nint(m), pos(m);
  DevBuf<uint8_t> sc2(scan_temp_bytes(m));
  const uint32_t root = 0;
  GWN_HIP(hipMemcpyAsync(fr[0].p, &root, 4, hipMemcpyHostToDevice, st));
  uint32_t F = 1, base = 0;
  int cur = 0;
  while (F > 0) {
    hipLaunchKernelGGL(k_expand, dim3(div_up(F, 256)), dim3(256), 0, st, fr[cur].p, (int)F, L.p, R.p, nbox.p,
                       members.p, nint.p);
    check_launch("k_expand");
    exclusive_scan_u32(nint.p, pos.p, F, total.p, sc2.p, st);
    uint32_t next = 0;
    GWN_HIP(hipMemcpyAsync(&next, total.p, 4, hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
    GWN_CHECK((int64_t)base + F + next <= (int64_t)m, GWN_ERR_INVALID, "lbvh: node count overflow");
    hipLaunchKernelGGL(k_emit, dim3(div_up(F, 256)), dim3(256), 0, st, (int)F, members.p, pos.p, base, base + F,
                       vals.p, tmp_nodes.p, fr[cur ^ 1].p);
    check_launch("k_emit");
    base += F;
    F = next;
    cur ^= 1;
  }
  const uint32_t nn = base;
  nodes.alloc(4 * (size_t)nn);
  GWN_HIP(hipMemcpyAsync(nodes.p, tmp_nodes.p, 16 * (size_t)nn, hipMemcpyDeviceToDevice, st));
  GWN_HIP(hipStreamSynchronize(st));
  (void)a256;
  return (int64_t)nn;
}

}  // namespace gwn
