// knn.hip -- k nearest neighbours (igl::knn, knn.cpp:24-127) and point areas
// (igl::copyleft::cgal::point_areas, point_areas.cpp:37-170) on gfx950: the
// pre-step that gives a point cloud its areas when the caller has none
// (copyleft/cgal/fast_winding_number.cpp:32-43).
//
// kNN.  The reference pops points and octree cells from one priority queue
// ordered by distance (cells by a lower bound of it), a leaf cell giving its
// first point only, until k points are out.  The answer is the k candidates
// nearest to the query in increasing distance, the candidates being the first
// point of every non-empty leaf.  Here:
//   tree   candidates sorted along a Morton curve and cut into buckets of 16;
//          a 4-wide LBVH over the bucket boxes (lbvh_build) laid out as a rope
//          of slot records: the slot's box (FP64 extents rounded outward to
//          float, so a box distance never exceeds the distance to a point in
//          it) and a skip link, or the bucket for a leaf slot
//   query  queries in Morton order.  A lane seeds its list with the k + 8
//          candidates either side of its own place on the curve (a binary
//          search of its code in the candidates' codes), so its bound is
//          finite and close from the start.  Then the wave walks the rope
//          coherently (one shared record index; each lane follows its own
//          depth-first branch and bound) and a lane enters a slot whose box is
//          not farther than its current k-th distance.  Candidates that beat
//          the k-th are parked and inserted in rounds (see Parked).
//          Candidates are ordered by (squared distance, index): equal
//          distances go to the lower index, so the result depends neither on
//          the visit order nor on the other queries of the wave.
// Squared distances are FP64 and summed ((dx dx + dy dy) + dz dz) as the
// reference's norm() sums a row.
//
// Point areas.  Per point (one lane): the neighbours I.row(i), of which those
// whose normal has a positive dot product with the first neighbour's normal
// are kept; with fewer than 3 left the area is 0.  A plane is fitted through
// them.  The reference takes the SVD of the mean-centred neighbours: T is the
// third right singular vector and the plane coordinates are the first two
// columns of U S.  Here the eigenvectors of the 3x3 scatter matrix (cyclic
// Jacobi in FP64) give the same directions, and the same coordinates up to a
// rotation or reflection within the plane, which changes no distance used
// below.  The reference then triangulates the projected neighbours (CGAL
// Delaunay) and sums, over the triangles at the first kept neighbour X, the
// area term of point_areas.cpp:115-160.  Only X's star is needed.  It is built
// by gift wrapping around X, starting from X's nearest neighbour (always a
// Delaunay edge).  The triangle left of edge X->a takes the point l, left of
// the edge, whose circle through X, a, l is met first as the circle's centre
// moves from the edge midpoint along the edge's left normal (the empty-circle
// condition).  The wrap goes counter-clockwise until it comes back to the
// first edge, or until no point lies left of the edge (X on the hull), and
// then clockwise from the first edge to the other side of the hull.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "gwn_internal.h"

namespace gwn {

constexpr int KNN_BUCKET = 16;             // candidates per leaf slot
constexpr uint32_t KN_LEAF = 0x80000000u;  // leaf slot: bucket index in the low bits
constexpr int KNN_MAX_REG = 64;            // larger k keeps its list in device memory
constexpr int PA_MAX_REG = 32;             // larger k keeps its plane points in device memory

namespace {

__device__ __forceinline__ float kn_round_down(double x) {
  const float f = (float)x;
  return ((double)f > x) ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float kn_round_up(double x) {
  const float f = (float)x;
  return ((double)f < x) ? nextafterf(f, INFINITY) : f;
}

// squared distance in the reference's summation order
__device__ __forceinline__ double kn_d2(const double* q, double x, double y, double z) {
  const double dx = x - q[0], dy = y - q[1], dz = z - q[2];
  return (dx * dx + dy * dy) + dz * dz;
}

__device__ __forceinline__ double kn_box_d2(const double* q, const float4& lo, const float4& hi) {
  // per axis: l - q below the box, q - h above it, 0 inside.  For a box
  // with l <= h at most one of the two differences is positive, so the max
  // picks the same value as the three-way branch (and 0 for a NaN q).
  const double l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
  double s = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double e = fmax(fmax(l[k] - q[k], q[k] - h[k]), 0.0);
    s += e * e;
  }
  return s;
}

// (a, ia) after (b, ib) in (distance, index) order
__device__ __forceinline__ bool kn_after(double a, int32_t ia, double b, int32_t ib) {
  return a > b || (a == b && ia > ib);
}

// ---------------- tree build ----------------
// bad[0] counts non-finite coordinates, bad[1] candidate indices outside V
__global__ void k_kn_coords(const double* __restrict__ V, int64_t nv, const int32_t* __restrict__ cand, int64_t nc,
                            float* __restrict__ Cf, uint32_t* __restrict__ bad) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nc) return;
  const int64_t v = cand ? (int64_t)cand[j] : j;
  if (v < 0 || v >= nv) {
    atomicAdd(bad + 1, 1u);
    Cf[3 * j] = Cf[3 * j + 1] = Cf[3 * j + 2] = 0.0f;
    return;
  }
  for (int a = 0; a < 3; ++a) {
    const double x = V[3 * v + a];
    if (!isfinite(x)) atomicAdd(bad, 1u);
    Cf[3 * j + a] = (float)x;
  }
}

// per block: min/max of float coordinates (6 values), reduced on the host
__global__ __launch_bounds__(256) void k_kn_bbox(const float* __restrict__ C, int64_t n, float* __restrict__ part) {
  __shared__ float sh[6][256];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], C[3 * j + a]);
      hi[a] = fmaxf(hi[a], C[3 * j + a]);
    }
  for (int a = 0; a < 3; ++a) {
    sh[a][threadIdx.x] = lo[a];
    sh[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        sh[a][threadIdx.x] = fminf(sh[a][threadIdx.x], sh[a][threadIdx.x + s]);
        sh[3 + a][threadIdx.x] = fmaxf(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// candidates in curve order: x, y, z and the index into V
__global__ void k_kn_points(const double* __restrict__ V, const int32_t* __restrict__ cand,
                            const uint32_t* __restrict__ order, int64_t nc, double4* __restrict__ pts) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
  const int64_t j = order[s];
  const int64_t v = cand ? (int64_t)cand[j] : j;
  pts[s] = make_double4(V[3 * v], V[3 * v + 1], V[3 * v + 2], (double)v);
}

// bucket boxes in the LBVH input layout (xmin, xmax, ymin, ymax, zmin, zmax)
__global__ void k_kn_bucket_boxes(const double4* __restrict__ pts, int64_t nc, int64_t nb, float* __restrict__ boxes) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int64_t s0 = b * KNN_BUCKET, s1 = s0 + KNN_BUCKET < nc ? s0 + KNN_BUCKET : nc;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t s = s0; s < s1; ++s) {
    const double4 p = pts[s];
    const double c[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = c[a] < lo[a] ? c[a] : lo[a];
      hi[a] = c[a] > hi[a] ? c[a] : hi[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    boxes[6 * b + 2 * a] = kn_round_down(lo[a]);
    boxes[6 * b + 2 * a + 1] = kn_round_up(hi[a]);
  }
}

// one level, deepest first: node box = union of its children's boxes;
// nsub = number of slot records below the node
__global__ void k_kn_up(const int32_t* __restrict__ level, int64_t cnt, const uint32_t* __restrict__ nodes,
                        const float* __restrict__ bboxes, float* __restrict__ nbox, uint32_t* __restrict__ nsub) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int64_t node = level[t];
  float b[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
  uint32_t n = 0;
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * node + s];
    if (c == NODE_EMPTY) break;
    const float* cb;
    if (c & NODE_INTERNAL) {
      const uint32_t ci = c & ~NODE_INTERNAL;
      cb = nbox + 6 * (int64_t)ci;
      n += 1 + nsub[ci];
    } else {
      cb = bboxes + 6 * (int64_t)c;
      n += 1;
    }
    for (int a = 0; a < 3; ++a) {
      b[2 * a] = fminf(b[2 * a], cb[2 * a]);
      b[2 * a + 1] = fmaxf(b[2 * a + 1], cb[2 * a + 1]);
    }
  }
  for (int k = 0; k < 6; ++k) nbox[6 * node + k] = b[k];
  nsub[node] = n;
}

// one level, root first: position of each internal child's first slot
__global__ void k_kn_pos(const int32_t* __restrict__ level, int64_t cnt, const uint32_t* __restrict__ nodes,
                         const uint32_t* __restrict__ nsub, uint32_t* __restrict__ first) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int64_t node = level[t];
  uint32_t pos = first[node];
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * node + s];
    if (c == NODE_EMPTY) break;
    if (c & NODE_INTERNAL) {
      const uint32_t ci = c & ~NODE_INTERNAL;
      first[ci] = pos + 1;
      pos += 1 + nsub[ci];
    } else {
      pos += 1;
    }
  }
}

// slot record (2 x float4): lo.xyz + link, hi.xyz; link = slot after the
// subtree for an internal child, KN_LEAF | bucket for a leaf
__global__ void k_kn_rec(int64_t nn, const uint32_t* __restrict__ nodes, const float* __restrict__ nbox,
                         const float* __restrict__ bboxes, const uint32_t* __restrict__ nsub,
                         const uint32_t* __restrict__ first, float4* __restrict__ rec) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  uint32_t pos = first[node];
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = nodes[4 * node + s];
    if (c == NODE_EMPTY) break;
    const float* b;
    uint32_t link, next;
    if (c & NODE_INTERNAL) {
      const uint32_t ci = c & ~NODE_INTERNAL;
      b = nbox + 6 * (int64_t)ci;
      next = pos + 1 + nsub[ci];
      link = next;
    } else {
      b = bboxes + 6 * (int64_t)c;
      next = pos + 1;
      link = KN_LEAF | c;
    }
    rec[2 * (int64_t)pos] = make_float4(b[0], b[2], b[4], __uint_as_float(link));
    rec[2 * (int64_t)pos + 1] = make_float4(b[1], b[3], b[5], 0.0f);
    pos = next;
  }
}

// ---------------- query ----------------
__global__ void k_kn_narrow(const double* __restrict__ Q, int64_t n, float* __restrict__ Qf,
                            uint32_t* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n) return;
  const double x = Q[i];
  if (!isfinite(x)) atomicAdd(bad, 1u);
  Qf[i] = (float)x;
}

// The k best so far, sorted descending in (distance, index); entry 0 is the
// current k-th.  Unused entries (j >= k) hold (-inf, -1) and never move.
template <int KM>
struct RegList {
  double d[KM];
  int32_t id[KM];
  __device__ __forceinline__ void init(int k, int64_t) {
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      d[j] = j < k ? INFINITY : -INFINITY;
      id[j] = j < k ? INT_MAX : -1;
    }
  }
  __device__ __forceinline__ double bound() const { return d[0]; }
  __device__ __forceinline__ bool better(double c, int32_t ci) const { return kn_after(d[0], id[0], c, ci); }
  // drop entry 0 and insert (c, ci) where it belongs
  __device__ __forceinline__ void insert(double c, int32_t ci) {
#pragma unroll
    for (int j = 0; j < KM - 1; ++j) {
      const bool nxt = kn_after(d[j + 1], id[j + 1], c, ci);
      const bool cur = kn_after(d[j], id[j], c, ci);
      const double nd = nxt ? d[j + 1] : (cur ? c : d[j]);
      const int32_t ni = nxt ? id[j + 1] : (cur ? ci : id[j]);
      d[j] = nd;
      id[j] = ni;
    }
    if (kn_after(d[KM - 1], id[KM - 1], c, ci)) {
      d[KM - 1] = c;
      id[KM - 1] = ci;
    }
  }
  __device__ __forceinline__ void write(int32_t* out, int k) const {
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (j < k) out[k - 1 - j] = id[j];
  }
};

// The same list in device memory (entry j of lane t at j * stride + t), for k
// above what registers hold.
struct MemList {
  double* d;
  int32_t* id;
  int64_t stride;
  int k;
  __device__ __forceinline__ void init(int kk, int64_t) {
    k = kk;
    for (int j = 0; j < k; ++j) {
      d[j * stride] = INFINITY;
      id[j * stride] = INT_MAX;
    }
  }
  __device__ __forceinline__ double bound() const { return d[0]; }
  __device__ __forceinline__ bool better(double c, int32_t ci) const { return kn_after(d[0], id[0], c, ci); }
  __device__ __forceinline__ void insert(double c, int32_t ci) {
    int j = 0;
    for (; j + 1 < k && kn_after(d[(j + 1) * stride], id[(j + 1) * stride], c, ci); ++j) {
      d[j * stride] = d[(j + 1) * stride];
      id[j * stride] = id[(j + 1) * stride];
    }
    d[j * stride] = c;
    id[j * stride] = ci;
  }
  __device__ __forceinline__ void write(int32_t* out, int kk) const {
    for (int j = 0; j < kk; ++j) out[kk - 1 - j] = id[j * stride];
  }
};

template <typename List>
__device__ __forceinline__ void kn_consider(List& L, const double* q, const double4& p) {
  const double d2 = kn_d2(q, p.x, p.y, p.z);
  const int32_t ci = (int32_t)p.w;
  if (L.better(d2, ci)) L.insert(d2, ci);
}

// Candidates that beat a lane's k-th are parked here and inserted in rounds:
// a wave then runs the sorted insert about once per round for all its lanes
// instead of once for every candidate that any one lane takes.
constexpr int KN_PARK = 4;
struct Parked {
  double d[KN_PARK];
  int32_t id[KN_PARK];
  int n = 0;
  __device__ __forceinline__ void put(double c, int32_t ci) {
#pragma unroll
    for (int t = 0; t < KN_PARK; ++t)
      if (t == n) {
        d[t] = c;
        id[t] = ci;
      }
    ++n;
  }
  template <typename List>
  __device__ __forceinline__ void flush(List& L) {
#pragma unroll
    for (int t = 0; t < KN_PARK; ++t)
      if (t < n && L.better(d[t], id[t])) L.insert(d[t], id[t]);
    n = 0;
  }
};

// First position in the sorted codes not below `code` (nc if none).
__device__ __forceinline__ int64_t kn_lower_bound(const uint32_t* __restrict__ codes, int64_t nc, uint32_t code) {
  int64_t lo = 0, hi = nc;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (codes[mid] < code)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Query t of the Morton-ordered batch (qcodes: its Morton code in the tree's
// box).  KM > 0: list in registers; KM == 0: list in device memory at md/mi
// (lane stride nq).
template <int KM>
__global__ __launch_bounds__(256) void k_knn(const float4* __restrict__ rec, uint32_t nrec,
                                             const double4* __restrict__ pts, const uint32_t* __restrict__ ccodes,
                                             int64_t nc, const double* __restrict__ Q, const uint32_t* __restrict__ perm,
                                             const uint32_t* __restrict__ qcodes, int64_t q0, int64_t nq, int k,
                                             double* __restrict__ md, int32_t* __restrict__ mi,
                                             int32_t* __restrict__ I) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const int64_t qi = perm[q0 + t];
  const double q[3] = {Q[3 * qi], Q[3 * qi + 1], Q[3 * qi + 2]};
  typename std::conditional<(KM > 0), RegList<(KM > 0 ? KM : 1)>, MemList>::type L;
  if constexpr (KM == 0) {
    L.d = md + t;
    L.id = mi + t;
    L.stride = nq;
  }
  L.init(k, t);
  // seed: the candidates around the query's own place on the Morton curve
  const int64_t c0 = kn_lower_bound(ccodes, nc, qcodes[q0 + t]);
  const int64_t half = k + KNN_BUCKET / 2;
  const int64_t w0 = c0 - half > 0 ? c0 - half : 0;
  const int64_t w1 = c0 + half < nc ? c0 + half : nc;
  for (int64_t s = w0; s < w1; ++s) kn_consider(L, q, pts[s]);
  // depth-first branch and bound over the rope, wave-coherent
  Parked park;
  uint32_t skip = 0, i = 0;
  while (i < nrec) {
    const float4 lo = rec[2 * (size_t)i], hi = rec[2 * (size_t)i + 1];
    const uint32_t link = __float_as_uint(lo.w);
    const bool leaf = (link & KN_LEAF) != 0;
    if (skip == i) {
      skip = i + 1;
      if (!(kn_box_d2(q, lo, hi) <= L.bound())) {
        if (!leaf) skip = link;
      } else if (leaf) {
        const int64_t s0 = (int64_t)(link & ~KN_LEAF) * KNN_BUCKET;
        const int64_t s1 = s0 + KNN_BUCKET < nc ? s0 + KNN_BUCKET : nc;
        for (int64_t s = s0; s < s1; ++s) {
          if (s < w0 || s >= w1) {
            const double4 p = pts[s];
            const double d2 = kn_d2(q, p.x, p.y, p.z);
            const int32_t ci = (int32_t)p.w;
            if (L.better(d2, ci)) park.put(d2, ci);
          }
          if (__ballot(park.n == KN_PARK) != 0) park.flush(L);
        }
      }
    }
    if (__ballot(park.n > 0) != 0) park.flush(L);
    const bool step = leaf || __ballot(skip == i + 1) != 0;
    i = __builtin_amdgcn_readfirstlane(step ? i + 1 : link);
  }
  L.write(I + qi * (int64_t)k, k);
}

// ---------------- point areas ----------------
template <int KM>
struct RegPlane {
  double x[KM], y[KM];
  __device__ __forceinline__ double& X(int j) { return x[j]; }
  __device__ __forceinline__ double& Y(int j) { return y[j]; }
};
struct MemPlane {
  double* x;
  double* y;
  int64_t stride;
  __device__ __forceinline__ double& X(int j) { return x[j * stride]; }
  __device__ __forceinline__ double& Y(int j) { return y[j * stride]; }
};

// cyclic Jacobi rotation of the symmetric 3x3 a in the (p, q) plane;
// columns p, q of v rotate with it
template <int p, int q>
__device__ __forceinline__ void jacobi_rotate(double (&a)[3][3], double (&v)[3][3]) {
  if (a[p][q] == 0.0) return;
  const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
  const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
  const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
  constexpr int r = 3 - p - q;
  const double apq = a[p][q];
  const double app = a[p][p] - t * apq, aqq = a[q][q] + t * apq;
  const double arp = c * a[r][p] - s * a[r][q], arq = s * a[r][p] + c * a[r][q];
  a[p][p] = app;
  a[q][q] = aqq;
  a[p][q] = a[q][p] = 0.0;
  a[r][p] = a[p][r] = arp;
  a[r][q] = a[q][r] = arq;
  for (int k = 0; k < 3; ++k) {
    const double vp = v[k][p], vq = v[k][q];
    v[k][p] = c * vp - s * vq;
    v[k][q] = s * vp + c * vq;
  }
}

// area term of triangle (X, Y, Z), counter-clockwise, at X
// (point_areas.cpp:121-155)
__device__ double star_term(double xx, double xy, double yx, double yy, double zx, double zy) {
  const double ex0 = yx - zx, ex1 = yy - zy, ey0 = xx - zx, ey1 = xy - zy, ez0 = yx - xx, ez1 = yy - xy;
  const double x = sqrt(ex0 * ex0 + ex1 * ex1);
  const double y = sqrt(ey0 * ey0 + ey1 * ey1);
  const double z = sqrt(ez0 * ez0 + ez1 * ez1);
  const double cosX = (z * z + y * y - x * x) / (2 * y * z);
  const double cosY = (z * z + x * x - y * y) / (2 * x * z);
  const double cosZ = (x * x + y * y - z * z) / (2 * y * x);
  double b0 = x * cosX, b1 = y * cosY, b2 = z * cosZ;
  const double bs = b0 + b1 + b2;
  b0 /= bs;
  b1 /= bs;
  b2 /= bs;
  const double full = 0.25 * sqrt((x + (y + z)) * (z - (x - y)) * (z + (x - y)) * (x + (y - z)));
  if (cosX < 0) return 0.5 * full;
  if (cosY < 0 || cosZ < 0) return 0.25 * full;
  return (b1 * full + b2 * full) / 2;
}

constexpr double kCollinear2 = 1e-20;  // (1e-10)^2

// One point.  LIM = KM for the register form (loops unrolled, compile-time
// indices), k for the memory form.
template <int KM, typename Plane>
__device__ void point_area(const double* __restrict__ P, int64_t n, const int32_t* __restrict__ Irow, int k,
                           const double* __restrict__ N, int64_t i, Plane& u, double* __restrict__ Aout,
                           double* __restrict__ Tout, uint32_t* __restrict__ bad) {
  constexpr bool REG = KM > 0;
  constexpr int UNR = REG ? KM : 1;  // full unroll in the register form only
  const int LIM = REG ? KM : k;
  Aout[i] = 0.0;
  if (Tout)
    for (int a = 0; a < 3; ++a) Tout[3 * i + a] = 0.0;
  for (int j = 0; j < k; ++j)
    if (Irow[j] < 0 || Irow[j] >= n) {  // an index outside P: reported, nothing read
      atomicAdd(bad, 1u);
      return;
    }
  // normal filter (point_areas.cpp:72-82): keep neighbours whose normal has a
  // positive dot product with the first neighbour's
  uint64_t keepm = 0;  // register form: bit j
  int m = 0, first = -1;
  double n0[3] = {0, 0, 0};
  const bool filter = N != nullptr && k > 1;
  if (filter) {
    const int64_t v0 = Irow[0];
    n0[0] = N[3 * v0];
    n0[1] = N[3 * v0 + 1];
    n0[2] = N[3 * v0 + 2];
  }
  double mean[3] = {0, 0, 0};
#pragma unroll UNR
  for (int j = 0; j < LIM; ++j) {
    if (j < k) {
      const int64_t v = Irow[j];
      bool keep = true;
      if (filter) keep = (n0[0] * N[3 * v] + n0[1] * N[3 * v + 1]) + n0[2] * N[3 * v + 2] > 0;
      if (REG) {
        if (keep) keepm |= 1ull << j;
      } else {
        u.Y(j) = keep ? 1.0 : 0.0;  // memory form: the flag rides in Y until the projection
      }
      if (keep) {
        if (first < 0) first = j;
        ++m;
        for (int a = 0; a < 3; ++a) mean[a] += P[3 * v + a];
      }
    }
  }
  if (m <= 2) return;
  for (int a = 0; a < 3; ++a) mean[a] /= m;
  // scatter matrix of the centred neighbours
  double c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll UNR
  for (int j = 0; j < LIM; ++j) {
    const bool keep = REG ? ((keepm >> j) & 1) != 0 : (j < k && u.Y(j) != 0.0);
    if (j < k && keep) {
      const int64_t v = Irow[j];
      const double e[3] = {P[3 * v] - mean[0], P[3 * v + 1] - mean[1], P[3 * v + 2] - mean[2]};
      for (int r = 0; r < 3; ++r)
        for (int s = r; s < 3; ++s) c[r][s] += e[r] * e[s];
    }
  }
  c[1][0] = c[0][1];
  c[2][0] = c[0][2];
  c[2][1] = c[1][2];
  double ev[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = c[0][1] * c[0][1] + c[0][2] * c[0][2] + c[1][2] * c[1][2];
    const double dia = c[0][0] * c[0][0] + c[1][1] * c[1][1] + c[2][2] * c[2][2];
    if (!(off > 1e-36 * dia)) break;
    jacobi_rotate<0, 1>(c, ev);
    jacobi_rotate<0, 2>(c, ev);
    jacobi_rotate<1, 2>(c, ev);
  }
  // eigenvalues descending: e0, e1 span the plane, e2 is its normal
  int o0 = 0, o1 = 1, o2 = 2;
  if (c[o0][o0] < c[o1][o1]) { const int t = o0; o0 = o1; o1 = t; }
  if (c[o1][o1] < c[o2][o2]) { const int t = o1; o1 = o2; o2 = t; }
  if (c[o0][o0] < c[o1][o1]) { const int t = o0; o0 = o1; o1 = t; }
  double e0[3], e1[3], tn[3];
  for (int a = 0; a < 3; ++a) {
    e0[a] = ev[a][o0];
    e1[a] = ev[a][o1];
    tn[a] = ev[a][o2];
  }
  if (N != nullptr && Tout) {  // point_areas.cpp:100-103: T faces the point's own normal
    const double dn = tn[0] * N[3 * i] + tn[1] * N[3 * i + 1] + tn[2] * N[3 * i + 2];
    if (dn < 0)
      for (int a = 0; a < 3; ++a) tn[a] = -tn[a];
  }
  if (Tout)
    for (int a = 0; a < 3; ++a) Tout[3 * i + a] = tn[a];
  // plane coordinates; dropped neighbours get NaN and take no part below
  double px = 0.0, py = 0.0;
#pragma unroll UNR
  for (int j = 0; j < LIM; ++j) {
    if (j < k) {
      const bool keep = REG ? ((keepm >> j) & 1) != 0 : u.Y(j) != 0.0;
      if (keep) {
        const int64_t v = Irow[j];
        const double e[3] = {P[3 * v] - mean[0], P[3 * v + 1] - mean[1], P[3 * v + 2] - mean[2]};
        const double ux = e[0] * e0[0] + e[1] * e0[1] + e[2] * e0[2];
        const double uy = e[0] * e1[0] + e[1] * e1[1] + e[2] * e1[2];
        u.X(j) = ux;
        u.Y(j) = uy;
        if (j == first) {
          px = ux;
          py = uy;
        }
      } else {
        u.X(j) = u.Y(j) = __longlong_as_double(0x7FF8000000000000ll);
      }
    } else if (REG) {
      u.X(j) = u.Y(j) = __longlong_as_double(0x7FF8000000000000ll);
    }
  }
  // X's nearest neighbour: a Delaunay edge (lowest index among equals)
  int j1 = -1;
  double best = INFINITY, ax = 0, ay = 0;
#pragma unroll UNR
  for (int j = 0; j < LIM; ++j) {
    if (j < k && j != first) {
      const double rx = u.X(j) - px, ry = u.Y(j) - py;
      const double r2 = rx * rx + ry * ry;
      if (r2 > 0.0 && r2 < best) {  // NaN (dropped) fails both
        best = r2;
        j1 = j;
        ax = rx;
        ay = ry;
      }
    }
  }
  if (j1 < 0) return;  // every kept neighbour coincides with X: no triangle
  double acc = 0.0;
  // side +1: counter-clockwise from X->j1; side -1: clockwise from X->j1
  for (int side = 1; side >= -1; side -= 2) {
    double cx = ax, cy = ay;  // current edge X->cur
    int cur = j1;
    bool closed = false;
    for (int step = 0; step <= LIM; ++step) {
      int nl = -1;
      double nt = INFINITY, nx = 0, ny = 0;
      const double cc = cx * cx + cy * cy;
#pragma unroll UNR
      for (int l = 0; l < LIM; ++l) {
        if (l < k && l != first && l != cur) {
          const double rx = u.X(l) - px, ry = u.Y(l) - py;
          const double cr = side * (cx * ry - cy * rx);
          // on this side of the edge, by more than rounding: a point within
          // |sin| <= 1e-10 of the edge's line is collinear (a hull edge
          // through collinear points makes no triangle; NaN fails too)
          if (cr > 0.0 && cr * cr > kCollinear2 * cc * (rx * rx + ry * ry)) {
            const double tl = ((rx * rx + ry * ry) - (cx * rx + cy * ry)) / (2.0 * cr);
            // equal circles: the point first met turning from the edge
            const bool take = nl < 0 || tl < nt || (tl == nt && side * (nx * ry - ny * rx) < 0.0);
            if (take) {
              nl = l;
              nt = tl;
              nx = rx;
              ny = ry;
            }
          }
        }
      }
      if (nl < 0) break;  // X is on the hull on this side
      // triangle (X, cur, nl) is counter-clockwise for side +1
      if (side > 0)
        acc += star_term(px, py, px + cx, py + cy, px + nx, py + ny);
      else
        acc += star_term(px, py, px + nx, py + ny, px + cx, py + cy);
      cur = nl;
      cx = nx;
      cy = ny;
      if (nl == j1) {
        closed = true;
        break;
      }
    }
    if (closed) break;  // X is interior: the counter-clockwise wrap saw every triangle
  }
  Aout[i] = isfinite(acc) ? acc : 0.0;
}

// points order[i0 .. i0 + m) of the n (Morton order, so a wavefront's
// neighbourhoods overlap in cache); the memory form keeps lane t's plane
// points at mx/my[j * m + t]
template <int KM>
__global__ __launch_bounds__(256) void k_point_areas(const double* __restrict__ P, int64_t n,
                                                     const uint32_t* __restrict__ order, int64_t i0, int64_t m,
                                                     const int32_t* __restrict__ I, int k,
                                                     const double* __restrict__ N, double* __restrict__ mx,
                                                     double* __restrict__ my, double* __restrict__ A,
                                                     double* __restrict__ T, uint32_t* __restrict__ bad) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int64_t i = order[i0 + t];
  if constexpr (KM > 0) {
    RegPlane<KM> u;
    point_area<KM>(P, n, I + i * (int64_t)k, k, N, i, u, A, T, bad);
  } else {
    MemPlane u{mx + t, my + t, m};
    point_area<0>(P, n, I + i * (int64_t)k, k, N, i, u, A, T, bad);
  }
}

// ---------------- exact duplicates ----------------
// FP64 bit patterns with -0 folded into +0 (equal doubles, equal keys)
__global__ void k_dup_key(const double* __restrict__ V, const uint32_t* __restrict__ order, int64_t n, int axis,
                          uint64_t* __restrict__ keys) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t v = order ? (int64_t)order[s] : s;
  keys[s] = (uint64_t)__double_as_longlong(V[3 * v + axis] + 0.0);
}

__global__ void k_to_float(const double* __restrict__ x, int64_t n, float* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

__global__ void k_iota(uint32_t* __restrict__ v, int64_t n) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) v[s] = (uint32_t)s;
}

// sorted by coordinates (stably, so equal points keep ascending indices):
// a point equal to its predecessor is not a candidate
__global__ void k_dup_flags(const double* __restrict__ V, const uint32_t* __restrict__ order, int64_t n,
                            uint32_t* __restrict__ keep) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  bool k = true;
  if (s > 0) {
    const int64_t a = order[s], b = order[s - 1];
    k = !(V[3 * a] == V[3 * b] && V[3 * a + 1] == V[3 * b + 1] && V[3 * a + 2] == V[3 * b + 2]);
  }
  keep[s] = k ? 1u : 0u;
}

__global__ void k_dup_compact(const uint32_t* __restrict__ order, const uint32_t* __restrict__ keep,
                              const uint32_t* __restrict__ pos, int64_t n, int32_t* __restrict__ cand) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && keep[s]) cand[pos[s]] = (int32_t)order[s];
}

}  // namespace

// ---------------- host side ----------------
struct KnnTree {
  int64_t nc = 0;
  uint32_t nrec = 0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};
  DevBuf<double4> pts;    // candidates in curve order
  DevBuf<uint32_t> codes; // their Morton codes, ascending
  DevBuf<float4> rec;
};

static void knn_bbox(const float* Cf, int64_t n, float* bb, hipStream_t st) {
  const int g = (int)std::min<int64_t>(1024, std::max<int64_t>(1, div_up(n, 256)));
  DevBuf<float> part(6 * (size_t)g);
  hipLaunchKernelGGL(k_kn_bbox, dim3(g), dim3(256), 0, st, Cf, n, part.p);
  check_launch("k_kn_bbox");
  std::vector<float> h(6 * (size_t)g);
  GWN_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  for (int a = 0; a < 3; ++a) {
    bb[a] = INFINITY;
    bb[3 + a] = -INFINITY;
  }
  for (int b = 0; b < g; ++b)
    for (int a = 0; a < 3; ++a) {
      bb[a] = std::min(bb[a], h[6 * b + a]);
      bb[3 + a] = std::max(bb[3 + a], h[6 * b + 3 + a]);
    }
}

static void knn_build(const double* V, int64_t nv, const int32_t* cand, int64_t nc, KnnTree& T, hipStream_t st) {
  T.nc = nc;
  DevBuf<float> Cf(3 * (size_t)nc);
  DevBuf<uint32_t> bad(2);
  GWN_HIP(hipMemsetAsync(bad.p, 0, 8, st));
  hipLaunchKernelGGL(k_kn_coords, dim3(div_up(nc, 256)), dim3(256), 0, st, V, nv, cand, nc, Cf.p, bad.p);
  check_launch("k_kn_coords");
  uint32_t nbad[2] = {0, 0};
  GWN_HIP(hipMemcpyAsync(nbad, bad.p, 8, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  GWN_CHECK(nbad[1] == 0, GWN_ERR_INVALID, "knn: a candidate index lies outside V");
  GWN_CHECK(nbad[0] == 0, GWN_ERR_INVALID, "knn: points must be finite");
  knn_bbox(Cf.p, nc, T.bbox, st);
  DevBuf<uint32_t> order(nc);
  T.codes.alloc((size_t)nc);
  morton_keys(Cf.p, nc, T.bbox, T.codes.p, order.p, st);
  {
    const size_t tb = radix_sort_temp_bytes(nc, false);
    DevBuf<uint8_t> tmp(tb);
    radix_sort_u32(T.codes.p, order.p, nc, 30, tmp.p, tb, st);
  }
  T.pts.alloc((size_t)nc);
  hipLaunchKernelGGL(k_kn_points, dim3(div_up(nc, 256)), dim3(256), 0, st, V, cand, order.p, nc, T.pts.p);
  check_launch("k_kn_points");
  const int64_t nb = div_up(nc, KNN_BUCKET);
  DevBuf<float> bboxes(6 * (size_t)nb);
  hipLaunchKernelGGL(k_kn_bucket_boxes, dim3(div_up(nb, 256)), dim3(256), 0, st, T.pts.p, nc, nb, bboxes.p);
  check_launch("k_kn_bucket_boxes");
  DevBuf<uint32_t> nodes;
  const int64_t nn = lbvh_build(bboxes.p, nb, T.bbox, nodes, st);
  GWN_CHECK(nn > 0, GWN_ERR_INVALID, "knn: empty tree");
  Levels L;
  compute_levels(nodes.p, nn, L, st);
  DevBuf<float> nbox(6 * (size_t)nn);
  DevBuf<uint32_t> nsub(nn), first(nn);
  for (int d = L.maxdepth; d >= 0; --d) {
    const int64_t cnt = L.offset[d + 1] - L.offset[d];
    if (!cnt) continue;
    hipLaunchKernelGGL(k_kn_up, dim3(div_up(cnt, 128)), dim3(128), 0, st, L.order.p + L.offset[d], cnt, nodes.p,
                       bboxes.p, nbox.p, nsub.p);
    check_launch("k_kn_up");
  }
  GWN_HIP(hipMemsetAsync(first.p, 0, 4, st));
  for (int d = 0; d <= L.maxdepth; ++d) {
    const int64_t cnt = L.offset[d + 1] - L.offset[d];
    if (!cnt) continue;
    hipLaunchKernelGGL(k_kn_pos, dim3(div_up(cnt, 256)), dim3(256), 0, st, L.order.p + L.offset[d], cnt, nodes.p,
                       nsub.p, first.p);
    check_launch("k_kn_pos");
  }
  uint32_t nrec = 0;
  GWN_HIP(hipMemcpyAsync(&nrec, nsub.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  T.nrec = nrec;
  T.rec.alloc(2 * (size_t)nrec);
  hipLaunchKernelGGL(k_kn_rec, dim3(div_up(nn, 256)), dim3(256), 0, st, nn, nodes.p, nbox.p, bboxes.p, nsub.p,
                     first.p, T.rec.p);
  check_launch("k_kn_rec");
  GWN_HIP(hipStreamSynchronize(st));
}

static void knn_query(const KnnTree& T, const double* Q, int64_t nq, int k, int32_t* I, hipStream_t st) {
  DevBuf<float> Qf(3 * (size_t)nq);
  DevBuf<uint32_t> bad(1), keys(nq), perm(nq);
  GWN_HIP(hipMemsetAsync(bad.p, 0, 4, st));
  hipLaunchKernelGGL(k_kn_narrow, dim3(div_up(3 * nq, 256)), dim3(256), 0, st, Q, nq, Qf.p, bad.p);
  check_launch("k_kn_narrow");
  uint32_t nbad = 0;
  GWN_HIP(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  GWN_CHECK(nbad == 0, GWN_ERR_INVALID, "knn: query points must be finite");
  morton_keys(Qf.p, nq, T.bbox, keys.p, perm.p, st);
  {
    const size_t tb = radix_sort_temp_bytes(nq, false);
    DevBuf<uint8_t> tmp(tb);
    radix_sort_u32(keys.p, perm.p, nq, 30, tmp.p, tb, st);
  }
  ProfileScope prof("knn", st);
#define GWN_KNN(KMV)                                                                                           \
  hipLaunchKernelGGL(k_knn<KMV>, dim3(div_up(nq, 256)), dim3(256), 0, st, T.rec.p, T.nrec, T.pts.p, T.codes.p, \
                     T.nc, Q, perm.p, keys.p, (int64_t)0, nq, k, (double*)nullptr, (int32_t*)nullptr, I)
  if (k <= 8)
    GWN_KNN(8);
  else if (k <= 16)
    GWN_KNN(16);
  else if (k <= 24)
    GWN_KNN(24);
  else if (k <= 32)
    GWN_KNN(32);
  else if (k <= 48)
    GWN_KNN(48);
  else if (k <= KNN_MAX_REG)
    GWN_KNN(64);
  else {
    // lists in device memory, in chunks of queries
    const int64_t chunk = std::max<int64_t>(256, std::min<int64_t>(nq, ((int64_t)1 << 28) / (12 * (int64_t)k)));
    DevBuf<double> md((size_t)chunk * k);
    DevBuf<int32_t> mi((size_t)chunk * k);
    for (int64_t q0 = 0; q0 < nq; q0 += chunk) {
      const int64_t m = std::min(chunk, nq - q0);
      hipLaunchKernelGGL(k_knn<0>, dim3(div_up(m, 256)), dim3(256), 0, st, T.rec.p, T.nrec, T.pts.p, T.codes.p,
                         T.nc, Q, perm.p, keys.p, q0, m, k, md.p, mi.p, I);
    }
  }
#undef GWN_KNN
  check_launch("k_knn");
}

int knn_columns(int64_t nv, int64_t k) { return (int)(nv <= k ? nv : k); }

// knn.cpp:24-127.  V nv x 3, P np x 3 (device, FP64); cand (device) lists the
// candidates (NULL: all of V); I np x knn_columns(nv, k) (device).
void knn_device(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand,
                int64_t ncand, int32_t* I, hipStream_t st) {
  GWN_CHECK(k >= 1 && k < (1 << 20), GWN_ERR_INVALID, "knn: k must be in [1, 2^20)");
  if (np == 0 || nv == 0) return;
  if (nv <= k) {
    // knn.cpp:46-55: every point of V, nearest first
    cand = nullptr;
    ncand = nv;
    k = nv;
  } else if (!cand) {
    ncand = nv;
  }
  GWN_CHECK(ncand >= k, GWN_ERR_INVALID, "knn: fewer candidate points (non-empty octree leaves) than k");
  KnnTree T;
  knn_build(V, nv, cand, ncand, T, st);
  knn_query(T, P, np, (int)k, I, st);
  GWN_HIP(hipStreamSynchronize(st));
}

// point_areas.cpp:37-170.  P n x 3, I n x k, N n x 3 or NULL, A n, T n x 3 or
// NULL; all device.
void point_areas_device(const double* P, int64_t n, const int32_t* I, int64_t k, const double* N, double* A,
                        double* T, hipStream_t st) {
  GWN_CHECK(k >= 1 && k < (1 << 20), GWN_ERR_INVALID, "point_areas: I must have 1 to 2^20 columns");
  if (n == 0) return;
  DevBuf<uint32_t> bad(1), keys(n), order(n);
  {
    DevBuf<float> Pf(3 * (size_t)n);
    GWN_HIP(hipMemsetAsync(bad.p, 0, 4, st));
    hipLaunchKernelGGL(k_kn_narrow, dim3(div_up(3 * n, 256)), dim3(256), 0, st, P, n, Pf.p, bad.p);
    check_launch("k_kn_narrow");
    float bb[6];
    knn_bbox(Pf.p, n, bb, st);
    morton_keys(Pf.p, n, bb, keys.p, order.p, st);
    const size_t tb = radix_sort_temp_bytes(n, false);
    DevBuf<uint8_t> tmp(tb);
    radix_sort_u32(keys.p, order.p, n, 30, tmp.p, tb, st);
    uint32_t nbad = 0;
    GWN_HIP(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
    GWN_CHECK(nbad == 0, GWN_ERR_INVALID, "point_areas: points must be finite");
  }
  ProfileScope prof("point_areas", st);
#define GWN_PA(KMV)                                                                                               \
  hipLaunchKernelGGL(k_point_areas<KMV>, dim3(div_up(n, 256)), dim3(256), 0, st, P, n, order.p, (int64_t)0, n, I, \
                     (int)k, N, (double*)nullptr, (double*)nullptr, A, T, bad.p)
  if (k <= 8)
    GWN_PA(8);
  else if (k <= 16)
    GWN_PA(16);
  else if (k <= 24)
    GWN_PA(24);
  else if (k <= PA_MAX_REG)
    GWN_PA(32);
  else {
    const int64_t chunk = std::max<int64_t>(256, std::min<int64_t>(n, ((int64_t)1 << 27) / (16 * k)));
    DevBuf<double> mx((size_t)chunk * k), my((size_t)chunk * k);
    for (int64_t i0 = 0; i0 < n; i0 += chunk) {
      const int64_t m = std::min(chunk, n - i0);
      hipLaunchKernelGGL(k_point_areas<0>, dim3(div_up(m, 256)), dim3(256), 0, st, P, n, order.p, i0, m, I, (int)k,
                         N, mx.p, my.p, A, T, bad.p);
    }
  }
#undef GWN_PA
  check_launch("k_point_areas");
  uint32_t nbad = 0;
  GWN_HIP(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  GWN_CHECK(nbad == 0, GWN_ERR_INVALID, "point_areas: an index of I lies outside P");
}

void to_float_device(const double* x, int64_t n, float* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_to_float, dim3(div_up(n, 256)), dim3(256), 0, st, x, n, y);
  check_launch("k_to_float");
}

// The candidates of knn over P's own octree: every point except exact copies
// of a lower-indexed point (igl::octree keeps coincident points in one leaf,
// of which knn.cpp:113-117 takes the first).  Returns the count.
int64_t distinct_points_device(const double* V, int64_t n, DevBuf<int32_t>& cand, hipStream_t st) {
  DevBuf<uint64_t> keys(n);
  DevBuf<uint32_t> order(n);
  hipLaunchKernelGGL(k_iota, dim3(div_up(n, 256)), dim3(256), 0, st, order.p, n);
  check_launch("k_iota");
  const size_t tb = radix_sort_temp_bytes(n, true);
  DevBuf<uint8_t> tmp(tb);
  for (int axis = 2; axis >= 0; --axis) {  // least significant axis first; stable passes
    hipLaunchKernelGGL(k_dup_key, dim3(div_up(n, 256)), dim3(256), 0, st, V, order.p, n, axis, keys.p);
    check_launch("k_dup_key");
    radix_sort_u64(keys.p, order.p, n, 64, tmp.p, tb, st);
  }
  DevBuf<uint32_t> keep(n), pos(n), total(1);
  hipLaunchKernelGGL(k_dup_flags, dim3(div_up(n, 256)), dim3(256), 0, st, V, order.p, n, keep.p);
  check_launch("k_dup_flags");
  DevBuf<uint8_t> sc(scan_temp_bytes(n));
  exclusive_scan_u32(keep.p, pos.p, n, total.p, sc.p, st);
  uint32_t m = 0;
  GWN_HIP(hipMemcpyAsync(&m, total.p, 4, hipMemcpyDeviceToHost, st));
  GWN_HIP(hipStreamSynchronize(st));
  cand.alloc(m);
  hipLaunchKernelGGL(k_dup_compact, dim3(div_up(n, 256)), dim3(256), 0, st, order.p, keep.p, pos.p, n, cand.p);
  check_launch("k_dup_compact");
  GWN_HIP(hipStreamSynchronize(st));
  return (int64_t)m;
}

}  // namespace gwn
