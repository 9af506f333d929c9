// antipodal.hip -- generalized winding numbers by the antipodal method:
// igl::WindingNumberAntipodalScene (WindingNumberAntipodalScene.h:48-256) and
// igl::winding_number_antipodal (winding_number_antipodal.cpp), with the
// ray-crossing count the reference asks of Embree
// (embree/EmbreeIntersector.cpp:390-443) computed here on the GPU.
//
//   W(p) = sum_e w_e * halfangle(x1, a_e - p, b_e - p) / (2 pi)  +  c(p)
//
// over the weighted open-boundary edges e (interior edges of an oriented
// surface cancel), x1 = -x0, and c(p) = the signed number of triangles the
// ray p + t x0 (t >= 0) crosses, +1 where x0 leaves through a triangle's
// front (x0 . (v1-v0)x(v2-v0) > 0), -1 otherwise.
//
//   boundary   the 3F directed edges as (min, max) keys with +-1, radix
//              sorted; each run of equal keys sums to the edge's net count;
//              non-zero nets become segments oriented so the weight is > 0
//   frac       per query, the boundary sum in FP64 (segments staged in LDS)
//   crossings  all rays share the direction x0, so the queries are taken in
//              Hilbert order and a wavefront walks the soup's rope of slot
//              boxes coherently (as the closest-point search does): a lane
//              enters a slot whose box its ray meets (FP64 slab test on boxes
//              rounded outward) and counts its triangle with the watertight
//              ray-triangle test of Woop, Benthin and Wald (JCGT 2013) in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gwn_internal.h"

namespace gwn {
namespace {

constexpr uint32_t AP_LEAF = 0x80000000u;  // rope leaf tag (the soup's)
constexpr double AP_PI = 3.1415926535897932384626433832795;

// ---------------- boundary ----------------
__global__ void k_ap_edges(const int32_t* __restrict__ F, int64_t nf, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ vals, int8_t* __restrict__ sgn) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= 3 * nf) return;
  const int64_t t = j / 3;
  const int e = (int)(j % 3);
  const uint32_t u = (uint32_t)F[3 * t + e], v = (uint32_t)F[3 * t + (e + 1) % 3];
  const uint32_t lo = u < v ? u : v, hi = u < v ? v : u;
  keys[j] = ((uint64_t)lo << 32) | hi;
  vals[j] = (uint32_t)j;
  sgn[j] = (int8_t)(u < v ? 1 : -1);
}

// net count of each run of equal keys, at the run's first position (0 elsewhere)
__global__ void k_ap_net(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                         const int8_t* __restrict__ sgn, int64_t n, int32_t* __restrict__ net,
                         uint32_t* __restrict__ flag) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int32_t sum = 0;
  if (s == 0 || keys[s] != keys[s - 1])
    for (int64_t t = s; t < n && keys[t] == keys[s]; ++t) sum += sgn[vals[t]];
  net[s] = sum;
  flag[s] = sum != 0 ? 1u : 0u;
}

// segment record: a.xyz, b.xyz, w, 0
__global__ void k_ap_segments(const uint64_t* __restrict__ keys, const int32_t* __restrict__ net,
                              const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int64_t n,
                              const double* __restrict__ V, double* __restrict__ seg) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !flag[s]) return;
  const int64_t lo = (int64_t)(keys[s] >> 32), hi = (int64_t)(keys[s] & 0xFFFFFFFFull);
  const int32_t w = net[s];
  const int64_t ia = w > 0 ? lo : hi, ib = w > 0 ? hi : lo;
  double* o = seg + 8 * (int64_t)pos[s];
  for (int k = 0; k < 3; ++k) {
    o[k] = V[3 * ia + k];
    o[3 + k] = V[3 * ib + k];
  }
  o[6] = (double)(w > 0 ? w : -w);
  o[7] = 0.0;
}

// ---------------- fractional term ----------------
// WindingNumberAntipodalScene.h:180-191: half the signed solid angle of the
// spherical triangle (x1, v0, v1), unnormalised Van Oosterom-Strackee
__device__ __forceinline__ double ap_dot(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
__device__ __forceinline__ double ap_half_angle(const double* x1, const double* v0, const double* v1) {
  const double l0 = sqrt(ap_dot(v0, v0)), l1 = sqrt(ap_dot(v1, v1));
  const double c[3] = {v0[1] * v1[2] - v0[2] * v1[1], v0[2] * v1[0] - v0[0] * v1[2], v0[0] * v1[1] - v0[1] * v1[0]};
  const double num = ap_dot(x1, c);
  const double den = ((l0 * l1 + l1 * ap_dot(x1, v0)) + l0 * ap_dot(x1, v1)) + ap_dot(v0, v1);
  return atan2(num, den);
}

constexpr int AP_TILE = 256;
__global__ __launch_bounds__(256) void k_ap_frac(const double* __restrict__ seg, int64_t nseg,
                                                 const double* __restrict__ O, int64_t no, double x1x, double x1y,
                                                 double x1z, double* __restrict__ W) {
  __shared__ double tile[AP_TILE * 7];
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = q < no;
  double p[3] = {0, 0, 0};
  if (live)
    for (int k = 0; k < 3; ++k) p[k] = O[3 * q + k];
  const double x1[3] = {x1x, x1y, x1z};
  double area = 0.0;
  for (int64_t base = 0; base < nseg; base += AP_TILE) {
    const int64_t m = nseg - base < AP_TILE ? nseg - base : AP_TILE;
    __syncthreads();
    for (int64_t t = threadIdx.x; t < m; t += blockDim.x)
      for (int k = 0; k < 7; ++k) tile[7 * t + k] = seg[8 * (base + t) + k];
    __syncthreads();
    if (live)
      for (int64_t t = 0; t < m; ++t) {
        const double* s = tile + 7 * t;
        const double v0[3] = {s[0] - p[0], s[1] - p[1], s[2] - p[2]};
        const double v1[3] = {s[3] - p[0], s[4] - p[1], s[5] - p[2]};
        area += s[6] * ap_half_angle(x1, v0, v1);
      }
  }
  if (live) W[q] = area / (2.0 * AP_PI);
}

__global__ void k_ap_zero(double* __restrict__ W, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) W[i] = 0.0;
}

__global__ void k_ap_narrow(const double* __restrict__ O, int64_t n, float* __restrict__ Of) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * n) Of[i] = (float)O[i];
}

// ---------------- signed crossings ----------------
struct Ray {
  double d[3];    // direction x0
  double inv[3];  // 1 / d (unused on an axis where d is 0)
  int kx, ky, kz; // watertight shear: kz the dominant axis
  double sx, sy, sz;
};

// does the ray o + t d, t >= 0, meet the box?  (boxes are rounded outward)
__device__ __forceinline__ bool ap_box(const Ray& r, const double* o, const float4& lo, const float4& hi) {
  const double l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
  double tmin = 0.0, tmax = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (r.d[a] == 0.0) {
      if (o[a] < l[a] || o[a] > h[a]) return false;
    } else {
      double t0 = (l[a] - o[a]) * r.inv[a], t1 = (h[a] - o[a]) * r.inv[a];
      if (t0 > t1) {
        const double t = t0;
        t0 = t1;
        t1 = t;
      }
      tmin = t0 > tmin ? t0 : tmin;
      tmax = t1 < tmax ? t1 : tmax;
    }
  }
  return tmin <= tmax;
}

// +1 / -1 for a crossing of triangle (A, B, C) at t >= 0, 0 for none
__device__ __forceinline__ int ap_cross(const Ray& r, const double* o, const double* A, const double* B,
                                        const double* C) {
  const double a[3] = {A[0] - o[0], A[1] - o[1], A[2] - o[2]};
  const double b[3] = {B[0] - o[0], B[1] - o[1], B[2] - o[2]};
  const double c[3] = {C[0] - o[0], C[1] - o[1], C[2] - o[2]};
  const double ax = a[r.kx] - r.sx * a[r.kz], ay = a[r.ky] - r.sy * a[r.kz];
  const double bx = b[r.kx] - r.sx * b[r.kz], by = b[r.ky] - r.sy * b[r.kz];
  const double cx = c[r.kx] - r.sx * c[r.kz], cy = c[r.ky] - r.sy * c[r.kz];
  const double U = cx * by - cy * bx, V = ax * cy - ay * cx, Wt = bx * ay - by * ax;
  if ((U < 0.0 || V < 0.0 || Wt < 0.0) && (U > 0.0 || V > 0.0 || Wt > 0.0)) return 0;
  const double det = U + V + Wt;
  if (det == 0.0) return 0;
  const double T = U * (r.sz * a[r.kz]) + V * (r.sz * b[r.kz]) + Wt * (r.sz * c[r.kz]);
  if (det > 0.0 ? T < 0.0 : T > 0.0) return 0;  // behind the origin
  const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
  const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  return (r.d[0] * n[0] + r.d[1] * n[1]) + r.d[2] * n[2] > 0.0 ? 1 : -1;
}

__global__ __launch_bounds__(256) void k_ap_crossings(const float4* __restrict__ dbox, uint32_t nslots,
                                                      const double* __restrict__ V, const int32_t* __restrict__ F,
                                                      const double* __restrict__ O, const uint32_t* __restrict__ perm,
                                                      int64_t no, Ray r, double* __restrict__ W) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= no) return;
  const int64_t qi = perm[t];
  const double o[3] = {O[3 * qi], O[3 * qi + 1], O[3 * qi + 2]};
  int count = 0;
  uint32_t skip = 0, i = 0;
  while (i < nslots) {
    const float4 lo = dbox[2 * (size_t)i], hi = dbox[2 * (size_t)i + 1];
    // the whole 32-byte record in one scalar-cache round trip, not the link
    // first and the box after it
    asm volatile("" ::"s"(lo.x), "s"(lo.w), "s"(hi.z));
    const uint32_t link = __float_as_uint(lo.w);
    const bool leaf = (link & AP_LEAF) != 0;
    if (skip == i) {
      skip = i + 1;
      if (!ap_box(r, o, lo, hi)) {
        if (!leaf) skip = link;
      } else if (leaf) {
        const int32_t* tf = F + 3 * (size_t)(link & ~AP_LEAF);
        const double A[3] = {V[3 * (int64_t)tf[0]], V[3 * (int64_t)tf[0] + 1], V[3 * (int64_t)tf[0] + 2]};
        const double B[3] = {V[3 * (int64_t)tf[1]], V[3 * (int64_t)tf[1] + 1], V[3 * (int64_t)tf[1] + 2]};
        const double C[3] = {V[3 * (int64_t)tf[2]], V[3 * (int64_t)tf[2] + 1], V[3 * (int64_t)tf[2] + 2]};
        count += ap_cross(r, o, A, B, C);
      }
    }
    const bool step = leaf || __ballot(skip == i + 1) != 0;
    i = __builtin_amdgcn_readfirstlane(step ? i + 1 : link);
  }
  W[qi] = W[qi] + (double)count;  // frac + c (WindingNumberAntipodalScene.h:104)
}

}  // namespace

void soup_build_into(gwn_soup* s, const double* V, int64_t nv, const int32_t* F, int64_t nf, int tree_kind);

// Boundary segments of (V, F) on the device (8 doubles each); returns the count.
static int64_t ap_boundary(const double* dV, const int32_t* dF, int64_t nf, DevBuf<double>& seg, hipStream_t st) {
  const int64_t n = 3 * nf;
  if (n == 0) return 0;
  DevBuf<uint64_t> keys(n);
  DevBuf<uint32_t> vals(n), flag(n), pos(n), total(1);
  DevBuf<int8_t> sgn(n);
  DevBuf<int32_t> net(n);
  hipLaunchKernelGGL(k_ap_edges, dim3(div_up(n, 256)), dim3(256), 0, st, dF, nf, keys.p, vals.p, sgn.p);
  check_launch("k_ap_edges");
  {
    const size_t tb = radix_sort_temp_bytes(n, true);
    DevBuf<uint8_t> tmp(tb);
    radix_sort_u64(keys.p, vals.p, n, 64, tmp.p, tb, st);
  }
  hipLaunchKernelGGL(k_ap_net, dim3(div_up(n, 256)), dim3(256), 0, st, keys.p, vals.p, sgn.p, n, net.p, flag.p);
  check_launch("k_ap_net");
  DevBuf<uint8_t> sc(scan_temp_bytes(n));
  exclusive_scan_u32(flag.p, pos.p, n, total.p, sc.p, st);
  uint32_t m = 0;
  GWN_HIP(hipMemcpyAsync(&m, total.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  seg.alloc(8 * (size_t)(m > 0 ? m : 1));
  if (m > 0) {
    hipLaunchKernelGGL(k_ap_segments, dim3(div_up(n, 256)), dim3(256), 0, st, keys.p, net.p, flag.p, pos.p, n, dV,
                       seg.p);
    check_launch("k_ap_segments");
  }
  GWN_HIP(hipStreamSynchronize(st));
  return (int64_t)m;
}

}  // namespace gwn

namespace gwn {

// WindingNumberAntipodalScene.h:69-78 (the scene) plus the ray structure
void antipodal_build(gwn_antipodal* a, const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* x0,
                     hipStream_t st) {
  if (x0) {
    for (int k = 0; k < 3; ++k) a->x0[k] = x0[k];
  } else {  // default_x0(): normalize(1, sqrt 2, sqrt 3) (WindingNumberAntipodalScene.h:167-173)
    const double d[3] = {1.0, std::sqrt(2.0), std::sqrt(3.0)};
    const double nrm = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) a->x0[k] = d[k] / nrm;
  }
  GWN_CHECK(std::isfinite(a->x0[0]) && std::isfinite(a->x0[1]) && std::isfinite(a->x0[2]) &&
                (a->x0[0] != 0.0 || a->x0[1] != 0.0 || a->x0[2] != 0.0),
            GWN_ERR_INVALID, "x0 must be a finite non-zero direction");
  soup_build_into(&a->soup, V, nv, F, nf, GWN_TREE_LBVH);
  if (nf > 0) soup_ensure_dist(&a->soup, st);
  a->nseg = ap_boundary(a->soup.V64.p, a->soup.F.p, nf, a->seg, st);
}

// WindingNumberAntipodalScene.h:82-150: O no x 3, W no (device, FP64)
void antipodal_query_device(const gwn_antipodal* a, const double* O, int64_t no, double* W, hipStream_t st) {
  if (no <= 0) return;
  const unsigned g = div_up(no, 256);
  if (a->nseg > 0) {
    ProfileScope prof("antipodal_frac", st);
    hipLaunchKernelGGL(k_ap_frac, dim3(g), dim3(256), 0, st, a->seg.p, a->nseg, O, no, -a->x0[0], -a->x0[1],
                       -a->x0[2], W);
    check_launch("k_ap_frac");
  } else {
    hipLaunchKernelGGL(k_ap_zero, dim3(g), dim3(256), 0, st, W, no);
    check_launch("k_ap_zero");
  }
  const gwn_soup* s = &a->soup;
  if (s->nf == 0 || s->nslots == 0) return;
  const size_t tb = radix_sort_temp_bytes(no, false);
  Scratch ws(scratch_bytes({12 * (size_t)no, 4 * (size_t)no, 4 * (size_t)no, tb}), st);
  float* Of = ws.take_n<float>(3 * (size_t)no);
  uint32_t* keys = ws.take_n<uint32_t>((size_t)no);
  uint32_t* perm = ws.take_n<uint32_t>((size_t)no);
  void* tmp = ws.take(tb);
  hipLaunchKernelGGL(k_ap_narrow, dim3(div_up(3 * no, 256)), dim3(256), 0, st, O, no, Of);
  check_launch("k_ap_narrow");
  hilbert_keys(Of, no, s->bbox, keys, perm, st);
  radix_sort_u32(keys, perm, no, 30, tmp, tb, st);
  Ray r;
  int kz = 0;
  for (int k = 1; k < 3; ++k)
    if (std::fabs(a->x0[k]) > std::fabs(a->x0[kz])) kz = k;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (a->x0[kz] < 0.0) std::swap(kx, ky);  // keeps the winding of the projected triangle
  for (int k = 0; k < 3; ++k) {
    r.d[k] = a->x0[k];
    r.inv[k] = a->x0[k] != 0.0 ? 1.0 / a->x0[k] : 0.0;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.sx = a->x0[kx] / a->x0[kz];
  r.sy = a->x0[ky] / a->x0[kz];
  r.sz = 1.0 / a->x0[kz];
  {
    ProfileScope prof("antipodal_crossings", st);
    hipLaunchKernelGGL(k_ap_crossings, dim3(g), dim3(256), 0, st, s->dbox.p, (uint32_t)s->nslots, s->V64.p, s->F.p,
                       O, perm, no, r, W);
    check_launch("k_ap_crossings");
  }
}

}  // namespace gwn
