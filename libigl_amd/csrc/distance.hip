// distance.hip -- closest points on a triangle soup and igl::signed_distance
// (signed_distance.cpp:28-140, 665-712) for the unsigned, winding-number and
// fast-winding-number sign types.
//
// The closest point uses the soup's own rope (the FWN tree) with a bounding
// box per slot record (soup_ensure_dist), walked wave-coherently like the
// winding-number traversal: each lane keeps the next record it needs and the
// wave steps through one shared index, so each lane visits the records of its
// own depth-first branch-and-bound walk in order and its result does not
// depend on which other queries share its wave.  Distances are FP64 from the
// FP64 vertices, as the reference computes them for double V.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "gwn_internal.h"
#include "ring_walk.h"

namespace gwn {
namespace {

constexpr uint32_t DL_LEAF = 0x80000000u;  // same tag as the rope links
// Records a k_closest wavefront steps through before it hands its queries to
// k_closest_ring: 1/64 of the tree, within [2048, 8192] (the best of the
// budgets tried on 81,920 and 1,310,720 triangles, 1M points each).
inline uint32_t closest_budget(int64_t nslots) {
  return (uint32_t)std::min<int64_t>(8192, std::max<int64_t>(2048, nslots / 64));
}

__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Closest point of triangle (A, B, C) to p by Voronoi regions (Ericson,
// Real-Time Collision Detection, 2005, sec. 5.1.5), with the region tests in
// the order and form of point_simplex_squared_distance.cpp:43-114 so that
// points on region boundaries resolve to the same feature.  Returns |p-c|^2.
__device__ double tri_closest(const double* p, const double* A, const double* B, const double* C, double* c) {
  double ab[3], ac[3], ap[3], bp[3], cp[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = B[k] - A[k];
    ac[k] = C[k] - A[k];
    ap[k] = p[k] - A[k];
  }
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  bool done = false;
  if (d1 <= 0.0 && d2 <= 0.0) {
    for (int k = 0; k < 3; ++k) c[k] = A[k];
    done = true;
  }
  double d3 = 0, d4 = 0;
  if (!done) {
    for (int k = 0; k < 3; ++k) bp[k] = p[k] - B[k];
    d3 = dot3(ab, bp);
    d4 = dot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
      for (int k = 0; k < 3; ++k) c[k] = B[k];
      done = true;
    }
  }
  double vc = 0;
  if (!done) {
    vc = d1 * d4 - d3 * d2;
    const bool ab_distinct = A[0] != B[0] || A[1] != B[1] || A[2] != B[2];
    if (ab_distinct && vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
      const double v = d1 / (d1 - d3);
      for (int k = 0; k < 3; ++k) c[k] = A[k] + v * ab[k];
      done = true;
    }
  }
  double d5 = 0, d6 = 0;
  if (!done) {
    for (int k = 0; k < 3; ++k) cp[k] = p[k] - C[k];
    d5 = dot3(ab, cp);
    d6 = dot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
      for (int k = 0; k < 3; ++k) c[k] = C[k];
      done = true;
    }
  }
  double vb = 0;
  if (!done) {
    vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
      const double w = d2 / (d2 - d6);
      for (int k = 0; k < 3; ++k) c[k] = A[k] + w * ac[k];
      done = true;
    }
  }
  if (!done) {
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
      const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      for (int k = 0; k < 3; ++k) c[k] = B[k] + w * (C[k] - B[k]);
    } else {
      const double denom = 1.0 / (va + vb + vc);
      const double v = vb * denom, w = vc * denom;
      for (int k = 0; k < 3; ++k) c[k] = A[k] + ab[k] * v + ac[k] * w;
    }
  }
  const double e0 = p[0] - c[0], e1 = p[1] - c[1], e2 = p[2] - c[2];
  return e0 * e0 + e1 * e1 + e2 * e2;
}

// squared distance from p to a box (0 inside), AlignedBox::squaredExteriorDistance
__device__ __forceinline__ double box_d2(const double* p, const float4& lo, const float4& hi) {
  // per axis: l - p below the box, p - h above it, 0 inside.  For a box
  // with l <= h at most one of the two differences is positive, so the max
  // picks the same value as the three-way branch (and 0 for a NaN p).
  const double l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
  double s = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double e = fmax(fmax(l[k] - p[k], p[k] - h[k]), 0.0);
    s += e * e;
  }
  return s;
}

__global__ void k_narrow3(const double* __restrict__ P, int64_t n, float* __restrict__ Pf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * n) Pf[i] = (float)P[i];
}

// |p - c|^2 for the closest point c on face f
__device__ __forceinline__ double face_dist(const double* p, const double* __restrict__ V,
                                            const int32_t* __restrict__ F, uint32_t f, double* c) {
  const int32_t* tf = F + 3 * (size_t)f;
  double A[3], B[3], C[3];
  for (int k = 0; k < 3; ++k) {
    A[k] = V[3 * (int64_t)tf[0] + k];
    B[k] = V[3 * (int64_t)tf[1] + k];
    C[k] = V[3 * (int64_t)tf[2] + k];
  }
  return tri_closest(p, A, B, C, c);
}

// Face f replaces the best so far when strictly closer (AABB.cpp set_min).
__device__ __forceinline__ bool face_closer(const double* p, const double* __restrict__ V,
                                            const int32_t* __restrict__ F, uint32_t f, double& best, double* bc) {
  double c[3];
  const double d2 = face_dist(p, V, F, f, c);
  if (!(d2 < best)) return false;
  best = d2;
  for (int k = 0; k < 3; ++k) bc[k] = c[k];
  return true;
}

// Initial bound: descend greedily into the child whose box is nearest (at
// most 4 siblings per level) down to one triangle.  The walks then replace it
// only by strictly closer triangles, so the answer is still the minimum; among
// equally close faces the greedy one is kept.
__device__ __forceinline__ void greedy_bound(const float4* __restrict__ dbox, uint32_t nslots,
                                             const double* __restrict__ V, const int32_t* __restrict__ F,
                                             const double* p, double& best, int32_t& bf, double* bc) {
  uint32_t lo = 0, hi = nslots;
  for (int depth = 0; depth < 128 && lo < hi; ++depth) {
    uint32_t pick = 0xFFFFFFFFu, plink = 0;
    double pd = best;
    for (uint32_t k = lo; k < hi;) {
      const float4 blo = dbox[2 * (size_t)k], bhi = dbox[2 * (size_t)k + 1];
      const uint32_t l = __float_as_uint(blo.w);
      const double d = box_d2(p, blo, bhi);
      if (d < pd) {
        pd = d;
        pick = k;
        plink = l;
      }
      k = (l & DL_LEAF) ? k + 1 : l;
    }
    if (pick == 0xFFFFFFFFu) break;
    if (plink & DL_LEAF) {
      const uint32_t f = plink & ~DL_LEAF;
      if (face_closer(p, V, F, f, best, bc)) bf = (int32_t)f;
      break;
    }
    lo = pick + 1;
    hi = plink;
  }
}

// Nearest triangle: depth-first branch and bound over the rope, a subtree
// entered only when its box is closer than the best distance so far
// (AABB.cpp:1135-1215 looks into a child when its exterior distance is below
// the current minimum).  `best` starts at up2, as in the reference.
//
// A wavefront walks the union of its 64 queries' walks, and for queries deep
// inside a closed surface (far from it in every direction) that union covers
// most of the tree: such wavefronts ran 30x the median and set the length of
// the whole launch.  A wavefront that has stepped through `budget` records
// hands its queries over to k_closest_ring (their indices appended to
// `deferred`, the count in *ndeferred) and writes nothing.
__global__ __launch_bounds__(256) void k_closest(const float4* __restrict__ dbox, uint32_t nslots,
                                                 const double* __restrict__ V, const int32_t* __restrict__ F,
                                                 const double* __restrict__ P, const uint32_t* __restrict__ perm,
                                                 int64_t np, double up2, uint32_t budget,
                                                 uint32_t* __restrict__ deferred, uint32_t* __restrict__ ndeferred,
                                                 double* __restrict__ best_out, int32_t* __restrict__ face_out,
                                                 double* __restrict__ c_out) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= np) return;
  const int64_t pi = perm ? (int64_t)perm[t] : t;
  const double p[3] = {P[3 * pi], P[3 * pi + 1], P[3 * pi + 2]};
  double best = up2, bc[3] = {0.0, 0.0, 0.0};
  int32_t bf = -1;
  greedy_bound(dbox, nslots, V, F, p, best, bf, bc);
  uint32_t skip = 0, i = 0, steps = 0;
  while (i < nslots) {
    if (++steps > budget) {
      // the wave's live lanes are lanes 0..n-1 (t < np is a prefix of the wave)
      const uint32_t lane = threadIdx.x & 63, live = (uint32_t)__popcll(__ballot(1));
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(ndeferred, live);
      base = __shfl(base, 0, 64);
      deferred[base + lane] = (uint32_t)pi;
      return;
    }
    const float4 lo = dbox[2 * (size_t)i], hi = dbox[2 * (size_t)i + 1];
    // the whole 32-byte record in one scalar-cache round trip, not the link
    // first and the box after it
    asm volatile("" ::"s"(lo.x), "s"(lo.w), "s"(hi.z));
    const uint32_t link = __float_as_uint(lo.w);
    const bool leaf = (link & DL_LEAF) != 0;
    if (skip == i) {
      skip = i + 1;
      if (!(box_d2(p, lo, hi) < best)) {
        if (!leaf) skip = link;
      } else if (leaf) {
        if (face_closer(p, V, F, link & ~DL_LEAF, best, bc)) bf = (int32_t)(link & ~DL_LEAF);
      }
    }
    const bool step = leaf || __ballot(skip == i + 1) != 0;
    i = __builtin_amdgcn_readfirstlane(step ? i + 1 : link);
  }
  best_out[pi] = best;
  face_out[pi] = bf;
  for (int k = 0; k < 3; ++k) c_out[3 * pi + k] = bc[k];
}

// The queries k_closest handed over, one wavefront per query: the greedy
// bound, then the 64 lanes share the branch and bound through ring_walk,
// pruning against the closest distance any lane has found (kept in LDS).
// The answer is the sequential walk's: the greedy face when no face is
// strictly closer, else the first face in rope order at the minimum
// distance.  That face's boxes are no farther than the minimum, which no
// bound here goes below, so it is always reached; ties between lanes go to
// the smaller rope position.  The grid strides over the list, whose length
// is read on the device.
__global__ __launch_bounds__(256) void k_closest_ring(const float4* __restrict__ dbox, uint32_t nslots,
                                                      const double* __restrict__ V, const int32_t* __restrict__ F,
                                                      const double* __restrict__ P,
                                                      const uint32_t* __restrict__ deferred,
                                                      const uint32_t* __restrict__ ndeferred, double up2,
                                                      double* __restrict__ best_out,
                                                      int32_t* __restrict__ face_out, double* __restrict__ c_out) {
  __shared__ uint2 ring_all[4][kReevalRing];
  __shared__ unsigned long long bound_all[4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint2* ring = ring_all[wv];
  const uint32_t n = *ndeferred;
  for (uint32_t w = blockIdx.x * 4 + wv; w < n; w += gridDim.x * 4) {
    const int64_t pi = deferred[w];
    const double p[3] = {P[3 * pi], P[3 * pi + 1], P[3 * pi + 2]};
    double gbest = up2, gc[3] = {0.0, 0.0, 0.0};
    int32_t gf = -1;
    greedy_bound(dbox, nslots, V, F, p, gbest, gf, gc);
    // distances are >= 0 (or NaN, never stored), so their bit patterns order
    // as unsigned integers
    if (lane == 0) bound_all[wv] = (unsigned long long)__double_as_longlong(gbest);
    // this lane's candidate: the closest face it evaluated, the earlier in
    // rope order on a tie; none (cpos = ~0) unless strictly closer than up2
    double cbest = up2, cc[3] = {0.0, 0.0, 0.0};
    uint32_t cpos = 0xFFFFFFFFu, cf = 0;
    ring_walk(ring, lane, nslots, [&](uint32_t i, uint32_t& link) {
      const float4 lo = dbox[2 * (size_t)i], hi = dbox[2 * (size_t)i + 1];
      link = __float_as_uint(lo.w);
      const double bound = __longlong_as_double((long long)bound_all[wv]);
      // boxes at the bound are entered: a face at the minimum distance is
      // reached whatever bound another lane has set
      if (box_d2(p, lo, hi) > bound) return true;
      if (!(link & DL_LEAF)) return false;
      const uint32_t f = link & ~DL_LEAF;
      double c[3];
      const double d2 = face_dist(p, V, F, f, c);
      if (d2 < cbest || (d2 == cbest && cpos != 0xFFFFFFFFu && i < cpos)) {
        cbest = d2;
        cpos = i;
        cf = f;
        for (int k = 0; k < 3; ++k) cc[k] = c[k];
      }
      if (d2 < bound) atomicMin(&bound_all[wv], (unsigned long long)__double_as_longlong(d2));
      return true;
    });
    // (distance, rope position) minimum over the lanes
    double rd = cbest;
    uint32_t rp = cpos;
    for (int off = 32; off >= 1; off >>= 1) {
      const double od = __shfl_xor(rd, off, 64);
      const uint32_t op = __shfl_xor(rp, off, 64);
      if (od < rd || (od == rd && op < rp)) {
        rd = od;
        rp = op;
      }
    }
    if (gf >= 0 && !(rd < gbest)) {  // no face strictly closer than the greedy one
      if (lane == 0) {
        best_out[pi] = gbest;
        face_out[pi] = gf;
        for (int k = 0; k < 3; ++k) c_out[3 * pi + k] = gc[k];
      }
    } else if (rp != 0xFFFFFFFFu) {
      if (cpos == rp) {
        best_out[pi] = cbest;
        face_out[pi] = (int32_t)cf;
        for (int k = 0; k < 3; ++k) c_out[3 * pi + k] = cc[k];
      }
    } else if (lane == 0) {  // nothing closer than up2
      best_out[pi] = up2;
      face_out[pi] = -1;
      for (int k = 0; k < 3; ++k) c_out[3 * pi + k] = 0.0;
    }
  }
}

// pseudonormal_test.cpp:24-141: the normal at the closest point c of face f
// -- the face's, an edge's or a vertex's, by where c lies on the face.
// Faces of double area > 1e-4 (Kahan's Heron formula over the sorted edge
// lengths, doublearea.cpp:144-197) classify c by barycentric coordinates
// (barycentric_coordinates.cpp:88-100) <= 1e-12; smaller faces by distances
// to the corners and to the edges (project_to_line_segment.cpp) < 1e-12.
struct Pseudonormals {
  const double* V;   // nv x 3
  const int32_t* F;  // nf x 3
  const double* FN;  // nf x 3
  const double* VN;  // nv x 3
  const double* EN;  // nf x 3 x 3: edge opposite corner c of face f
};

__device__ __forceinline__ double len3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

__device__ void pseudonormal_at(const Pseudonormals& pn, int64_t f, const double* c, double* n) {
  constexpr double kMinDoubleArea = 1e-4, kEps = 1e-12;
  const int32_t* t = pn.F + 3 * f;
  const double* A = pn.V + 3 * (int64_t)t[0];
  const double* B = pn.V + 3 * (int64_t)t[1];
  const double* C = pn.V + 3 * (int64_t)t[2];
  // edge lengths |B-C|, |C-A|, |A-B| sorted descending
  double l0 = len3(B[0] - C[0], B[1] - C[1], B[2] - C[2]);
  double l1 = len3(C[0] - A[0], C[1] - A[1], C[2] - A[2]);
  double l2 = len3(A[0] - B[0], A[1] - B[1], A[2] - B[2]);
  double tmp;
  if (l0 < l1) tmp = l0, l0 = l1, l1 = tmp;
  if (l1 < l2) tmp = l1, l1 = l2, l2 = tmp;
  if (l0 < l1) tmp = l0, l0 = l1, l1 = tmp;
  const double arg = (l0 + (l1 + l2)) * (l2 - (l0 - l1)) * (l2 + (l0 - l1)) * (l0 + (l1 - l2));
  const double area = 2.0 * 0.25 * sqrt(arg);
  const double* src = pn.FN + 3 * f;
  if (area > kMinDoubleArea) {
    const double v0[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const double v1[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    const double v2[3] = {c[0] - A[0], c[1] - A[1], c[2] - A[2]};
    const double d00 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    const double d01 = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2];
    const double d11 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    const double d20 = v2[0] * v0[0] + v2[1] * v0[1] + v2[2] * v0[2];
    const double d21 = v2[0] * v1[0] + v2[1] * v1[1] + v2[2] * v1[2];
    const double denom = d00 * d11 - d01 * d01;
    double b[3];
    b[1] = (d11 * d20 - d01 * d21) / denom;
    b[2] = (d00 * d21 - d01 * d20) / denom;
    b[0] = 1.0 - (b[1] + b[2]);
    const int type = (b[0] <= kEps) + (b[1] <= kEps) + (b[2] <= kEps);
    if (type == 2) {  // a vertex
      for (int x = 0; x < 3; ++x)
        if (b[x] > kEps) {
          src = pn.VN + 3 * (int64_t)t[x];
          break;
        }
    } else if (type == 1) {  // an edge
      for (int x = 0; x < 3; ++x)
        if (b[x] <= kEps) {
          src = pn.EN + 9 * f + 3 * x;
          break;
        }
    }
  } else {
    bool found = false;
    for (int v = 0; v < 3 && !found; ++v) {
      const double* P = pn.V + 3 * (int64_t)t[v];
      if (len3(c[0] - P[0], c[1] - P[1], c[2] - P[2]) < kEps) {
        found = true;
        src = pn.VN + 3 * (int64_t)t[v];
      }
    }
    for (int e = 0; e < 3 && !found; ++e) {
      const double* S = pn.V + 3 * (int64_t)t[(e + 1) % 3];
      const double* D = pn.V + 3 * (int64_t)t[(e + 2) % 3];
      const double dms[3] = {D[0] - S[0], D[1] - S[1], D[2] - S[2]};
      const double vl = dms[0] * dms[0] + dms[1] * dms[1] + dms[2] * dms[2];
      const double smp[3] = {S[0] - c[0], S[1] - c[1], S[2] - c[2]};
      const double tt = -(dms[0] * smp[0] + dms[1] * smp[1] + dms[2] * smp[2]) / vl;
      double sq;
      if (tt < 0.0) {
        sq = (c[0] - S[0]) * (c[0] - S[0]) + (c[1] - S[1]) * (c[1] - S[1]) + (c[2] - S[2]) * (c[2] - S[2]);
      } else if (tt > 1.0) {
        sq = (c[0] - D[0]) * (c[0] - D[0]) + (c[1] - D[1]) * (c[1] - D[1]) + (c[2] - D[2]) * (c[2] - D[2]);
      } else {
        const double pp[3] = {(1.0 - tt) * S[0] + tt * D[0], (1.0 - tt) * S[1] + tt * D[1],
                              (1.0 - tt) * S[2] + tt * D[2]};
        sq = (c[0] - pp[0]) * (c[0] - pp[0]) + (c[1] - pp[1]) * (c[1] - pp[1]) + (c[2] - pp[2]) * (c[2] - pp[2]);
      }
      if (sqrt(sq) < kEps) {
        src = pn.EN + 9 * f + 3 * e;
        found = true;
      }
    }
  }
  n[0] = src[0];
  n[1] = src[1];
  n[2] = src[2];
}

// signed_distance.cpp:96-135: out of [low2, up2) -> NaN, F.rows()+1, 0;
// otherwise S = s * sqrt(d2) with s from the sign type.  N (pseudonormal sign
// only, when given) receives the normal the test used; 0 for points out of
// bounds (the reference leaves those rows unset).
__global__ void k_sd_combine(int64_t np, const double* __restrict__ P, const double* __restrict__ best,
                             const int32_t* __restrict__ face, const double* __restrict__ cl,
                             const double* __restrict__ W, int sign_type, double low2, double up2, int64_t nf,
                             Pseudonormals pn, double* __restrict__ S, int32_t* __restrict__ I,
                             double* __restrict__ C, double* __restrict__ N) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double d2 = best[p];
  if (d2 >= up2 || d2 < low2) {
    S[p] = __longlong_as_double(0x7FF8000000000000ll);
    I[p] = (int32_t)(nf + 1);
    for (int k = 0; k < 3; ++k) C[3 * p + k] = 0.0;
    if (N)
      for (int k = 0; k < 3; ++k) N[3 * p + k] = 0.0;
    return;
  }
  double s = 1.0;
  if (sign_type == GWN_SIGN_FAST_WINDING_NUMBER) {
    const double w = (double)(float)W[p];  // the reference's FWN value is a float
    s = 1.0 - 2.0 * fabs(w);
  } else if (sign_type == GWN_SIGN_WINDING_NUMBER || sign_type == GWN_SIGN_DEFAULT) {
    s = 1.0 - 2.0 * W[p];
  } else if (sign_type == GWN_SIGN_PSEUDONORMAL) {
    const double c[3] = {cl[3 * p], cl[3 * p + 1], cl[3 * p + 2]};
    double n[3];
    pseudonormal_at(pn, face[p], c, n);
    const double qc[3] = {P[3 * p] - c[0], P[3 * p + 1] - c[1], P[3 * p + 2] - c[2]};
    s = (qc[0] * n[0] + qc[1] * n[1] + qc[2] * n[2] >= 0.0) ? 1.0 : -1.0;
    if (N)
      for (int k = 0; k < 3; ++k) N[3 * p + k] = n[k];
  }
  S[p] = s * sqrt(d2);
  I[p] = face[p];
  for (int k = 0; k < 3; ++k) C[3 * p + k] = cl[3 * p + k];
}

__global__ void k_sd_empty(int64_t np, double* __restrict__ S, int32_t* __restrict__ I, double* __restrict__ C) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double nan = __longlong_as_double(0x7FF8000000000000ll);
  S[p] = nan;
  I[p] = -1;
  for (int k = 0; k < 3; ++k) C[3 * p + k] = nan;
}

}  // namespace

void soup_query_device(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision, double* W,
                       hipStream_t st);

// signed_distance.cpp:79-84: bounds on the unsigned squared distance.
static void sqr_bounds(double lower, double upper, double& low2, double& up2) {
  const auto mx = [](double a, double b) { return (a < b) ? b : a; };  // std::max
  const double max_abs = mx(std::fabs(lower), std::fabs(upper));
  up2 = std::pow(max_abs, 2.0);
  low2 = std::pow(mx(max_abs - (upper - lower), 0.0), 2.0);
}

// The pseudonormals of the soup's mesh (normals.cpp), built on first use from
// the vertices and faces as given.
static void soup_ensure_pseudonormals(gwn_soup* s, hipStream_t st) {
  if (s->pn_ready.load(std::memory_order_acquire)) return;
  std::lock_guard<std::mutex> lock(s->pn_mu);
  if (s->pn_ready.load(std::memory_order_relaxed)) return;
  std::vector<double> V(3 * (size_t)s->nv), FN, VN, EN;
  std::vector<int32_t> F(3 * (size_t)s->nf);
  GWN_HIP(hipMemcpyAsync(V.data(), s->V64.p, 8 * V.size(), hipMemcpyDeviceToHost, st));
  GWN_HIP(hipMemcpyAsync(F.data(), s->F.p, 4 * F.size(), hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  pseudonormals_host(V.data(), s->nv, F.data(), s->nf, FN, VN, EN);
  DevBuf<double> fn(FN.size()), vn(VN.size()), en(EN.size());
  GWN_HIP(hipMemcpyAsync(fn.p, FN.data(), 8 * FN.size(), hipMemcpyHostToDevice, st));
  GWN_HIP(hipMemcpyAsync(vn.p, VN.data(), 8 * VN.size(), hipMemcpyHostToDevice, st));
  GWN_HIP(hipMemcpyAsync(en.p, EN.data(), 8 * EN.size(), hipMemcpyHostToDevice, st));
  GWN_HIP(hipStreamSynchronize(st));
  s->pn_face = std::move(fn);
  s->pn_vertex = std::move(vn);
  s->pn_edge = std::move(en);
  s->pn_ready.store(true, std::memory_order_release);
}

void signed_distance_device(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower, double upper,
                            double* S, int32_t* I, double* C, double* N, hipStream_t st) {
  if (np <= 0) return;
  const unsigned g = div_up(np, 256);
  if (s->nf == 0) {  // signed_distance.cpp:397-403
    hipLaunchKernelGGL(k_sd_empty, dim3(g), dim3(256), 0, st, np, S, I, C);
    check_launch("k_sd_empty");
    return;
  }
  soup_ensure_dist(s, st);
  const bool pseudo = sign_type == GWN_SIGN_PSEUDONORMAL;
  if (pseudo) soup_ensure_pseudonormals(s, st);
  double low2, up2;
  sqr_bounds(lower, upper, low2, up2);
  const bool want_w = sign_type == GWN_SIGN_FAST_WINDING_NUMBER || sign_type == GWN_SIGN_WINDING_NUMBER ||
                      sign_type == GWN_SIGN_DEFAULT;
  const size_t tb = radix_sort_temp_bytes(np, false);
  const size_t n = (size_t)np;
  Scratch ws(scratch_bytes({12 * n, 4 * n, 4 * n, tb, 8 * n, 24 * n, 4 * n, want_w ? 8 * n : 8, 4}), st);
  float* Pf = ws.take_n<float>(3 * n);
  uint32_t* keys = ws.take_n<uint32_t>(n);
  uint32_t* perm = ws.take_n<uint32_t>(n);
  void* tmp = ws.take(tb);
  double* best = ws.take_n<double>(n);
  double* cl = ws.take_n<double>(3 * n);
  int32_t* face = ws.take_n<int32_t>(n);
  double* W = want_w ? ws.take_n<double>(n) : nullptr;
  uint32_t* ndeferred = ws.take_n<uint32_t>(1);
  hipLaunchKernelGGL(k_narrow3, dim3(div_up(3 * np, 256)), dim3(256), 0, st, P, np, Pf);
  check_launch("k_narrow3");
  hilbert_keys(Pf, np, s->bbox, keys, perm, st);
  radix_sort_u32(keys, perm, np, 30, tmp, tb, st);
  {
    ProfileScope prof("closest_point", st);
    // the sort keys are spent: their space holds the handed-over queries
    uint32_t* deferred = keys;
    fill_device<uint32_t>(ndeferred, 1, 0u, st);
    hipLaunchKernelGGL(k_closest, dim3(g), dim3(256), 0, st, s->dbox.p, (uint32_t)s->nslots, s->V64.p, s->F.p, P,
                       perm, np, up2, closest_budget(s->nslots), deferred, ndeferred, best, face, cl);
    check_launch("k_closest");
    const unsigned gr = (unsigned)std::min<int64_t>(div_up(np, 4), 2048);
    hipLaunchKernelGGL(k_closest_ring, dim3(gr), dim3(256), 0, st, s->dbox.p, (uint32_t)s->nslots, s->V64.p, s->F.p,
                       P, deferred, ndeferred, up2, best, face, cl);
    check_launch("k_closest_ring");
  }
  gwn_exact_tree hier;
  if (sign_type == GWN_SIGN_FAST_WINDING_NUMBER) {
    // signed_distance.cpp:125-130: fast_winding_number(fwn_bvh, 2, q.cast<float>())
    soup_query_device(s, Pf, np, 2.0f, GWN_PREC_FP32, W, st);
  } else if (sign_type == GWN_SIGN_WINDING_NUMBER || sign_type == GWN_SIGN_DEFAULT) {
    // signed_distance.cpp:60-64, 121-124: the exact winding number through a
    // WindingNumberAABB built for the call (hier3.set_mesh; hier3.grow())
    std::vector<double> V(3 * (size_t)s->nv);
    std::vector<int32_t> F(3 * (size_t)s->nf);
    GWN_HIP(hipMemcpyAsync(V.data(), s->V64.p, 8 * V.size(), hipMemcpyDeviceToHost, st));
    GWN_HIP(hipMemcpyAsync(F.data(), s->F.p, 4 * F.size(), hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
    exact_tree_build(&hier, V.data(), s->nv, F.data(), s->nf);
    wntree_query_device(hier.rec.p, hier.nrec, hier.tri.p, hier.bbox, P, np, W, st);
  }
  const Pseudonormals pn{s->V64.p, s->F.p, s->pn_face.p, s->pn_vertex.p, s->pn_edge.p};
  hipLaunchKernelGGL(k_sd_combine, dim3(g), dim3(256), 0, st, np, P, best, face, cl, W, sign_type, low2, up2, s->nf,
                     pn, S, I, C, pseudo ? N : nullptr);
  check_launch("k_sd_combine");
  if (sign_type == GWN_SIGN_WINDING_NUMBER || sign_type == GWN_SIGN_DEFAULT)
    GWN_HIP(hipStreamSynchronize(st));  // the hierarchy's records are freed on return
}

}  // namespace gwn
