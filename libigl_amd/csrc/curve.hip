// curve.hip -- 2D fast winding number of a curve (UT_SubtendedAngle) on gfx950.
//
// Build (gwn_curve_build):
//   k_seg_setup      per segment: its float box (initBounds/enlargeBounds,
//                    FastWindingNumberForSoups.h:7365-7390) and a packed
//                    a,b record for the exact angle
//   tree             host SAH4 over 2D boxes (the reference topology,
//                    BVH<4>::init<BOX_AREA,S,2>, 7392) or the device LBVH
//   k_curve_upward   one launch per tree level, deepest first: the
//                    reference's item()/post() moment formulas (7453-7599) ->
//                    BoxData, LocalData for the parent, subtree slot count
//   rope_first_slots top-down first slot of every node (soup.hip)
//   k_curve_slots    one 48-byte record per non-empty child slot
// Query (gwn_curve_query*): large batches are put in 2D Hilbert order (LDS radix
// sort); k_curve_traverse walks the rope layout, one query per lane: a far
// child slot (|q-P|^2 > maxP2*beta^2 and a finite Taylor value,
// AngleFunctors::pre 7725-7800) adds its order-0/1/2 term and jumps past its
// subtree, a near internal slot steps into its first child, a near segment
// adds the exact angle (UTsignedAngleSegment, 6279-6307).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_math.h"
#include "gwn_internal.h"

namespace gwn {
namespace {

constexpr uint32_t kLinkLeaf = 0x80000000u;
// batches of at least this many queries are put in Hilbert order first
constexpr int64_t kCurveSortMin = 262144;

__device__ __forceinline__ float cmin(float a, float b) { return (a > b) ? b : a; }
__device__ __forceinline__ float cmax(float a, float b) { return (a < b) ? b : a; }

__global__ void k_seg_setup(const float* __restrict__ P, const int32_t* __restrict__ E, int64_t ne,
                            float* __restrict__ boxes, float4* __restrict__ seg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const int64_t p0 = E[2 * i], p1 = E[2 * i + 1];
  const float ax = P[2 * p0], ay = P[2 * p0 + 1], bx = P[2 * p1], by = P[2 * p1 + 1];
  float* b = boxes + 6 * i;
  b[0] = cmin(ax, bx);
  b[1] = cmax(ax, bx);
  b[2] = cmin(ay, by);
  b[3] = cmax(ay, by);
  b[4] = 0.0f;
  b[5] = 0.0f;
  seg[i] = make_float4(ax, ay, bx, by);
}

// LocalData (7409-7427): what a child hands to its parent.
struct Local2 {
  float box[4];  // x0 x1 y0 y1
  float P[2], LP[2], N[2];
  float length;
  float Nd[2];
  float Nxy, Nyx;
  float Nkd[2];
  float C1, C2;  // 2Nxxy+Nyxx, 2Nyyx+Nxyy
};

// BoxData (7287-7320), 48 values per node: [0..3] maxP2, [4..11] P (x4, y4),
// [12..19] N, [20..27] NijDiag, [28..31] Nxy+Nyx, [32..39] NijkDiag,
// [40..43] 2Nxxy+Nyxx, [44..47] 2Nyyx+Nxyy.
enum : int { B2_MAXP2 = 0, B2_P = 4, B2_N = 12, B2_ND = 20, B2_NXY = 28, B2_NKD = 32, B2_C1 = 40, B2_C2 = 44 };

// PrecomputeFunctors::item (7453-7516).  The translation unit is compiled
// with -ffp-contract=off, so each expression rounds as the reference's.
__device__ void seg_local(const float4* __restrict__ seg, const float* __restrict__ boxes, uint32_t itemi, int order,
                          Local2& d) {
  const float4 s = seg[itemi];
  const float* bx = boxes + 6 * (int64_t)itemi;
  for (int k = 0; k < 4; ++k) d.box[k] = bx[k];
  const float ab0 = s.z - s.x, ab1 = s.w - s.y;
  const float N0 = ab1, N1 = -ab0;
  const float length2 = ab0 * ab0 + ab1 * ab1;
  const float length = sqrtf(length2);
  const float P0 = 0.5f * (s.x + s.z), P1 = 0.5f * (s.y + s.w);
  d.P[0] = P0;
  d.P[1] = P1;
  d.LP[0] = P0 * length;
  d.LP[1] = P1 * length;
  d.N[0] = N0;
  d.N[1] = N1;
  d.length = length;
  d.Nd[0] = d.Nd[1] = 0.0f;
  d.Nxy = d.Nyx = 0.0f;
  d.Nkd[0] = d.Nkd[1] = 0.0f;
  d.C1 = d.C2 = 0.0f;
  if (order < 2 || length == 0.0f) return;
  const float ixx = ab0 * ab0 / 12.0f;
  const float ixy = ab0 * ab1 / 12.0f;
  const float iyy = ab1 * ab1 / 12.0f;
  d.Nkd[0] = ixx * N0;
  d.Nkd[1] = iyy * N1;
  const float Nxxy = N0 * ixy, Nyxx = N1 * ixx, Nyyx = N1 * ixy, Nxyy = N0 * iyy;
  d.C1 = 2.0f * Nxxy + Nyxx;
  d.C2 = 2.0f * Nyyx + Nxyy;
}

// PrecomputeFunctors::post (7518-7599): combine the children into the node.
__device__ void node_post2(const Local2* ch, int nch, int order, float* bd, Local2& out) {
  float N[2] = {ch[0].N[0], ch[0].N[1]};
  float LP[2] = {ch[0].LP[0], ch[0].LP[1]};
  float length = ch[0].length;
  for (int k = 0; k < 2; ++k) {
    bd[B2_N + 4 * k] = N[k];
    bd[B2_P + 4 * k] = ch[0].P[k];
  }
  for (int i = 1; i < nch; ++i) {
    for (int k = 0; k < 2; ++k) {
      N[k] += ch[i].N[k];
      bd[B2_N + 4 * k + i] = ch[i].N[k];
      LP[k] += ch[i].LP[k];
      bd[B2_P + 4 * k + i] = ch[i].P[k];
    }
    length += ch[i].length;
  }
  for (int i = nch; i < 4; ++i)
    for (int k = 0; k < 2; ++k) bd[B2_N + 4 * k + i] = bd[B2_P + 4 * k + i] = 0.0f;
  out.N[0] = N[0];
  out.N[1] = N[1];
  out.LP[0] = LP[0];
  out.LP[1] = LP[1];
  out.length = length;
  float box[4];
  for (int k = 0; k < 4; ++k) box[k] = ch[0].box[k];
  for (int i = 1; i < nch; ++i)
    for (int a = 0; a < 2; ++a) {
      box[2 * a] = (box[2 * a] < ch[i].box[2 * a]) ? box[2 * a] : ch[i].box[2 * a];
      box[2 * a + 1] = (box[2 * a + 1] > ch[i].box[2 * a + 1]) ? box[2 * a + 1] : ch[i].box[2 * a + 1];
    }
  if (length > 0.0f) {
    const float r = 1.0f / length;  // UT_FixedVector::operator/ multiplies by the reciprocal
    out.P[0] = LP[0] * r;
    out.P[1] = LP[1] * r;
  } else {
    out.P[0] = 0.5f * (box[0] + box[1]);
    out.P[1] = 0.5f * (box[2] + box[3]);
  }
  for (int k = 0; k < 4; ++k) out.box[k] = box[k];
  for (int i = 0; i < nch; ++i) {
    const float d0 = cmax(ch[i].P[0] - ch[i].box[0], ch[i].box[1] - ch[i].P[0]);
    const float d1 = cmax(ch[i].P[1] - ch[i].box[2], ch[i].box[3] - ch[i].P[1]);
    bd[B2_MAXP2 + i] = d0 * d0 + d1 * d1;
  }
  for (int i = nch; i < 4; ++i) bd[B2_MAXP2 + i] = INFINITY;

  out.Nd[0] = ch[0].Nd[0];
  out.Nd[1] = ch[0].Nd[1];
  out.Nxy = out.Nyx = 0.0f;
  out.Nkd[0] = ch[0].Nkd[0];
  out.Nkd[1] = ch[0].Nkd[1];
  out.C1 = ch[0].C1;
  out.C2 = ch[0].C2;
  for (int i = 1; i < nch; ++i) {
    for (int k = 0; k < 2; ++k) {
      out.Nd[k] += ch[i].Nd[k];
      out.Nkd[k] += ch[i].Nkd[k];
    }
    out.C1 += ch[i].C1;
    out.C2 += ch[i].C2;
  }
  for (int i = 0; i < 4; ++i) {
    const bool on = i < nch;
    for (int k = 0; k < 2; ++k) {
      bd[B2_ND + 4 * k + i] = on ? ch[i].Nd[k] : 0.0f;
      bd[B2_NKD + 4 * k + i] = on ? ch[i].Nkd[k] : 0.0f;
    }
    bd[B2_NXY + i] = on ? ch[i].Nxy + ch[i].Nyx : 0.0f;
    bd[B2_C1 + i] = on ? ch[i].C1 : 0.0f;
    bd[B2_C2 + i] = on ? ch[i].C2 : 0.0f;
  }
  if (order < 1) return;
  for (int i = 0; i < nch; ++i) {
    const Local2& c = ch[i];
    const float dp[2] = {c.P[0] - out.P[0], c.P[1] - out.P[1]};
    const float* n = c.N;
    out.Nd[0] += n[0] * dp[0];
    out.Nd[1] += n[1] * dp[1];
    const float Nxy = c.Nxy + n[0] * dp[1];
    const float Nyx = c.Nyx + n[1] * dp[0];
    out.Nxy += Nxy;
    out.Nyx += Nyx;
    if (order >= 2) {
      for (int k = 0; k < 2; ++k) out.Nkd[k] += (2.0f * dp[k]) * c.Nd[k] + (dp[k] * dp[k]) * n[k];
      out.C1 += 2.0f * (dp[1] * c.Nd[0] + dp[0] * c.Nxy + n[0] * dp[0] * dp[1]) + 2.0f * c.Nyx * dp[0] +
                n[1] * dp[0] * dp[0];
      out.C2 += 2.0f * (dp[0] * c.Nd[1] + dp[1] * c.Nyx + n[1] * dp[1] * dp[0]) + 2.0f * c.Nxy * dp[1] +
                n[0] * dp[1] * dp[1];
    }
  }
}

__global__ void k_curve_upward(const int32_t* __restrict__ level_nodes, int64_t count,
                               const uint32_t* __restrict__ nodes, const float4* __restrict__ seg,
                               const float* __restrict__ boxes, int order, Local2* __restrict__ local,
                               float* __restrict__ boxdata, uint32_t* __restrict__ nslots_sub) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t node = level_nodes[t];
  Local2 ch[4];
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
      seg_local(seg, boxes, c, order, ch[s]);
      cnt += 1;
    }
    ++nch;
  }
  float bd[48];
  Local2 out;
  node_post2(ch, nch, order, bd, out);
  local[node] = out;
  for (int k = 0; k < 48; ++k) boxdata[48 * (int64_t)node + k] = bd[k];
  nslots_sub[node] = cnt;
}

// Slot record (3 x float4):
//   0: P.x P.y maxP2 link        1: -N.x -N.y A D        2: g30 g03 g21 g12
// The reference's far field (7737-7787), with x = q/|q| and m = 1/|q|,
//   -m (x.N) + m^2 (Ndx + Ndy - 2 (x^2 Ndx + y^2 Ndy + xy Nxy))
//     + m^3 (x.(3 Nkd + (C2, C1)) - 4 (x^3 Nkdx + y^3 Nkdy + x^2 y C1 + x y^2 C2)),
// is rewritten with x^2 + y^2 = 1 as m [ -(x.N) + m ( Qf(x) + m G(x) ) ] where
// Qf = A (x^2 - y^2) + D xy, A = Ndy - Ndx, D = -2 Nxy, and
// G = g30 x^3 + g03 y^3 + g21 x^2 y + g12 x y^2, g30 = C2 - Nkdx,
// g03 = C1 - Nkdy, g21 = 3 (Nkdy - C1), g12 = 3 (Nkdx - C2).  The
// coefficients are formed in FP64 from the reference's FP32 BoxData and
// rounded once.  link = skip index for an internal child, kLinkLeaf|segment
// for an item.
__global__ void k_curve_slots(int64_t nn, const uint32_t* __restrict__ nodes, const float* __restrict__ boxdata,
                              const uint32_t* __restrict__ nslots_sub, const uint32_t* __restrict__ first_slot,
                              float4* __restrict__ slots) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  const float* bd = boxdata + 48 * node;
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
      link = kLinkLeaf | c;
    }
    const double ndx = bd[B2_ND + s], ndy = bd[B2_ND + 4 + s], nxy = bd[B2_NXY + s];
    const double kx = bd[B2_NKD + s], ky = bd[B2_NKD + 4 + s], c1 = bd[B2_C1 + s], c2 = bd[B2_C2 + s];
    float4* r = slots + 3 * (int64_t)pos;
    r[0] = make_float4(bd[B2_P + s], bd[B2_P + 4 + s], bd[B2_MAXP2 + s], __uint_as_float(link));
    r[1] = make_float4(-bd[B2_N + s], -bd[B2_N + 4 + s], (float)(ndy - ndx), (float)(-2.0 * nxy));
    r[2] = make_float4((float)(c2 - kx), (float)(c1 - ky), (float)(3.0 * (ky - c1)), (float)(3.0 * (kx - c2)));
    pos = next;
  }
}

// Far-field value of one child slot from its record (see k_curve_slots), for
// q = query - P and l2 = |q|^2.
template <int ORDER>
__device__ __forceinline__ float far_field2(float x, float y, float l2, const float4& b, const float4& c) {
  const float m = __builtin_amdgcn_rsqf(l2);
  x *= m;
  y *= m;
  float t = fmaf(y, b.y, x * b.x);  // -(x.N)
  if (ORDER >= 1) {
    float qf = fmaf(b.w * x, y, b.z * ((x - y) * (x + y)));
    if (ORDER >= 2) {
      const float g = fmaf(x * x, fmaf(c.x, x, c.z * y), (y * y) * fmaf(c.w, x, c.y * y));
      qf = fmaf(m, g, qf);
    }
    t = fmaf(m, qf, t);
  }
  return m * t;
}

// UTsignedAngleSegment (6279-6307): the zero tests see the reference's
// rounding (no contraction in the numerator and denominator).
__device__ __forceinline__ float segment_angle(const float4 s, float qx, float qy) {
  const float ax = s.x - qx, ay = s.y - qy, bx = s.z - qx, by = s.w - qy;
  if ((ax == 0.0f && ay == 0.0f) || (bx == 0.0f && by == 0.0f)) return 0.0f;
  const float num = __fsub_rn(__fmul_rn(ax, by), __fmul_rn(ay, bx));
  if (num == 0.0f) return 0.0f;
  const float den = __fadd_rn(__fmul_rn(ax, bx), __fmul_rn(ay, by));
  return atan2_fast(num, den);
}

// Rope walk, one query per lane: each lane follows its own skip pointers and
// gathers its own 48-byte records (a far child slot adds its Taylor value and
// skips its subtree, a near internal slot steps into its first child, a near
// segment adds the exact angle).  A wave-coherent walk over a shared record
// index (as the soup's k_traverse) measured slower here at every batch size
// (1M queries on a 262,144-segment curve: 127 us against 115 us): the 2D
// records are small and a lane's path is short (~11 nodes), so the shared
// index mostly serialises lanes whose paths differ.
template <int ORDER>
__global__ __launch_bounds__(128) void k_curve_traverse(const float4* __restrict__ slots, uint32_t nslots,
                                                        const float4* __restrict__ seg, const float* __restrict__ Q,
                                                        const uint32_t* __restrict__ perm, int64_t nq, float acc2,
                                                        float* __restrict__ angle) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const int64_t qi = perm ? (int64_t)perm[t] : t;
  const float qx = Q[2 * qi], qy = Q[2 * qi + 1];
  float sum = 0.0f;
  uint32_t i = 0;
  while (i < nslots) {
    const float4* r = slots + 3 * (size_t)i;
    const float4 a = r[0], b = r[1];
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (ORDER >= 2) c = r[2];
    const uint32_t link = __float_as_uint(a.w);
    const bool leaf = (link & kLinkLeaf) != 0;
    const float x = qx - a.x, y = qy - a.y;
    const float l2 = __fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y));
    bool exact = leaf;
    uint32_t next = i + 1;
    if (!(l2 <= __fmul_rn(a.z, acc2))) {
      const float o = far_field2<ORDER>(x, y, l2, b, c);
      if (isfinite(o)) {
        sum = __fadd_rn(sum, o);
        exact = false;
        if (!leaf) next = link;
      }
    }
    if (exact) sum = __fadd_rn(sum, segment_angle(seg[link & ~kLinkLeaf], qx, qy));
    i = next;
  }
  angle[qi] = sum;
}

// 16-bit 2D Hilbert key of a query (a 256 x 256 grid over the curve's box;
// the classic rotate-and-reflect walk down the quadrants): two radix passes
// order a large batch well enough (wider keys cost more in the sort than they
// save in the walk)
__device__ __forceinline__ uint32_t hilbert2_8(uint32_t x, uint32_t y) {
  uint32_t d = 0;
  for (uint32_t s = 128; s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = 255u - x;
        y = 255u - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}
__global__ void k_order2(const float* __restrict__ Q, int64_t nq, float lx, float ly, float sx, float sy,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  float fx = (Q[2 * i] - lx) * sx, fy = (Q[2 * i + 1] - ly) * sy;
  fx = fminf(fmaxf(fx, 0.0f), 255.0f);
  fy = fminf(fmaxf(fy, 0.0f), 255.0f);
  keys[i] = hilbert2_8((uint32_t)fx, (uint32_t)fy);
  vals[i] = (uint32_t)i;
}

}  // namespace

void curve_build_device(gwn_curve* c, hipStream_t st) {
  const int64_t ne = c->ne;
  c->boxes.alloc(6 * (size_t)ne);
  c->seg.alloc((size_t)ne);
  if (ne == 0) {
    GWN_HIP(hipStreamSynchronize(st));
    return;
  }
  hipLaunchKernelGGL(k_seg_setup, dim3(div_up(ne, 256)), dim3(256), 0, st, c->P.p, c->E.p, ne, c->boxes.p, c->seg.p);
  check_launch("k_seg_setup");
  if (c->tree_kind == GWN_TREE_SAH4) {
    std::vector<float> boxes(6 * (size_t)ne);
    GWN_HIP(hipMemcpyAsync(boxes.data(), c->boxes.p, sizeof(float) * boxes.size(), hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
    std::vector<uint32_t> nodes;
    sah4_build(boxes.data(), (uint32_t)ne, nodes, 2);
    c->nnodes = (int64_t)(nodes.size() / 4);
    if (c->nnodes) {
      c->nodes.alloc(nodes.size());
      GWN_HIP(hipMemcpyAsync(c->nodes.p, nodes.data(), sizeof(uint32_t) * nodes.size(), hipMemcpyHostToDevice, st));
    }
  } else {
    const float bb3[6] = {c->bbox[0], c->bbox[1], 0.0f, c->bbox[2], c->bbox[3], 0.0f};
    c->nnodes = lbvh_build(c->boxes.p, ne, bb3, c->nodes, st);
  }
  const int64_t nn = c->nnodes;
  if (nn == 0) {
    c->nslots = 0;
    GWN_HIP(hipStreamSynchronize(st));
    return;
  }
  Levels L;
  compute_levels(c->nodes.p, nn, L, st);
  c->depth = L.maxdepth;
  c->boxdata.alloc(48 * (size_t)nn);
  DevBuf<uint32_t> nsub(nn), first(nn);
  {
    DevBuf<Local2> local(nn);
    for (int d = L.maxdepth; d >= 0; --d) {
      const int64_t cnt = L.offset[d + 1] - L.offset[d];
      if (!cnt) continue;
      hipLaunchKernelGGL(k_curve_upward, dim3(div_up(cnt, 128)), dim3(128), 0, st, L.order.p + L.offset[d], cnt,
                         c->nodes.p, c->seg.p, c->boxes.p, c->order, local.p, c->boxdata.p, nsub.p);
      check_launch("k_curve_upward");
    }
    GWN_HIP(hipStreamSynchronize(st));
  }
  uint32_t total = 0;
  GWN_HIP(hipMemcpyAsync(&total, nsub.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  rope_first_slots(L, c->nodes.p, nsub.p, first.p, st);
  GWN_HIP(hipStreamSynchronize(st));
  c->nslots = total;
  c->slots.alloc(3 * (size_t)total);
  hipLaunchKernelGGL(k_curve_slots, dim3(div_up(nn, 256)), dim3(256), 0, st, nn, c->nodes.p, c->boxdata.p, nsub.p,
                     first.p, c->slots.p);
  check_launch("k_curve_slots");
  GWN_HIP(hipStreamSynchronize(st));
}

void curve_query_device(const gwn_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle,
                        hipStream_t st) {
  if (nq <= 0) return;
  if (c->nslots == 0) {
    fill_device<float>(angle, nq, 0.0f, st);
    return;
  }
  // Large batches are walked in Hilbert order (neighbouring lanes then follow
  // similar paths); below kCurveSortMin the sort costs more than it saves
  // (65,536 queries: 32 us unsorted, 88 us sorted).
  uint32_t* perm = nullptr;
  const bool sorted = nq >= kCurveSortMin;
  const size_t tmpb = sorted ? radix_sort_temp_bytes(nq, false) : 0;
  // scratch of this call (sorted batches): keys, the permutation, sort temp
  Scratch ws(sorted ? scratch_bytes({4 * (size_t)nq, 4 * (size_t)nq, tmpb}) : 1, st);
  if (sorted) {
    uint32_t* keys = ws.take_n<uint32_t>((size_t)nq);
    perm = ws.take_n<uint32_t>((size_t)nq);
    void* tmp = ws.take(tmpb);
    const float ex = std::max(c->bbox[2] - c->bbox[0], 1e-30f), ey = std::max(c->bbox[3] - c->bbox[1], 1e-30f);
    hipLaunchKernelGGL(k_order2, dim3(div_up(nq, 256)), dim3(256), 0, st, Q, nq, c->bbox[0], c->bbox[1],
                       256.0f / ex, 256.0f / ey, keys, perm);
    check_launch("k_order2");
    ProfileScope prof("radix_sort", st);
    radix_sort_u32(keys, perm, nq, 16, tmp, tmpb, st);
  }
  const float acc2 = accuracy_scale * accuracy_scale;
  // blocks of 128 (1M queries: 131 us against 137 us with 256; 2 queries per
  // lane to overlap the two record chains measured 173-186 us)
  const unsigned g = div_up(nq, 128);
  const uint32_t ns = (uint32_t)c->nslots;
  ProfileScope prof("curve_traverse", st);
  switch (c->order) {
    case 0: hipLaunchKernelGGL(k_curve_traverse<0>, dim3(g), dim3(128), 0, st, c->slots.p, ns, c->seg.p, Q, perm, nq, acc2, angle); break;
    case 1: hipLaunchKernelGGL(k_curve_traverse<1>, dim3(g), dim3(128), 0, st, c->slots.p, ns, c->seg.p, Q, perm, nq, acc2, angle); break;
    default: hipLaunchKernelGGL(k_curve_traverse<2>, dim3(g), dim3(128), 0, st, c->slots.p, ns, c->seg.p, Q, perm, nq, acc2, angle); break;
  }
  check_launch("k_curve_traverse");
}

}  // namespace gwn
