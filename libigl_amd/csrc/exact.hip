// exact.hip -- exact FP64 generalized winding number.
//
// W(o) = sum over faces of solid_angle(f, O(o)) (winding_number.cpp:62-89,
// solid_angle.cpp:10-55), all in FP64.  The reference sums the same angles
// through its WindingNumberAABB hierarchy (cap triangles for boxes that do
// not contain the query); in real arithmetic both sums are equal.
//
// Kernel: one query per thread, 256 threads; triangles are staged through LDS
// in tiles of 256 (9 doubles each, 18 KB), so every triangle is read from
// HBM once per block and broadcast from LDS to all lanes.  A small query set
// also splits the triangles into chunks (grid.y) so that the GPU is full;
// the chunk partials are then added in chunk order.  FP64 VALU bound.  atan2
// terms are summed and divided by 2 pi once per query (the reference divides
// each term; the difference is a few ulps of the FP64 result).
#include <hip/hip_runtime.h>

#include "gwn_internal.h"
#include "ring_walk.h"
#include "wntree.h"

namespace gwn {

constexpr int EX_THREADS = 256;
constexpr int EX_TILE = 256;
constexpr int EX_QPT = 1;
constexpr double EX_PI = 3.1415926535897932384626433832795;

__device__ __forceinline__ double vos_atan2(const double* t, double px, double py, double pz) {
  const double a0 = t[0] - px, a1 = t[1] - py, a2 = t[2] - pz;
  const double b0 = t[3] - px, b1 = t[4] - py, b2 = t[5] - pz;
  const double c0 = t[6] - px, c1 = t[7] - py, c2 = t[8] - pz;
  const double la = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
  const double lb = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
  const double lc = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  // det[a;b;c] in the reference's term order
  const double det = a0 * b1 * c2 + b0 * c1 * a2 + c0 * a1 * b2 - c0 * b1 * a2 - b0 * a1 * c2 - a0 * c1 * b2;
  const double dbc = b0 * c0 + b1 * c1 + b2 * c2;
  const double dca = c0 * a0 + c1 * a1 + c2 * a2;
  const double dab = a0 * b0 + a1 * b1 + a2 * b2;
  return atan2(det, la * lb * lc + dbc * la + dca * lb + dab * lc);
}

__global__ __launch_bounds__(EX_THREADS) void k_exact_gather(const double* __restrict__ V,
                                                               const int32_t* __restrict__ F, int64_t nf,
                                                               double* __restrict__ T) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) T[9 * f + 3 * k + a] = V[3 * (int64_t)F[3 * f + k] + a];
}

// Block (bx, by): queries [bx * EX_THREADS * EX_QPT, ...) against the
// triangles of chunk by, [by * chunk, (by + 1) * chunk); the partial sum of
// query q over chunk by goes to part[by * no + q] (or to W / 2 pi directly
// when there is one chunk).  Splitting the triangles gives a small query set
// enough blocks to fill the GPU; the partials are added in chunk order.
__global__ __launch_bounds__(EX_THREADS) void k_exact(const double* __restrict__ T, int64_t nf, int64_t chunk,
                                                       const double* __restrict__ O, int64_t no,
                                                       double* __restrict__ part, double* __restrict__ W) {
  __shared__ double tile[EX_TILE * 9];
  const int64_t q0 = ((int64_t)blockIdx.x * EX_THREADS) * EX_QPT + threadIdx.x;
  double px[EX_QPT], py[EX_QPT], pz[EX_QPT], acc[EX_QPT];
  for (int k = 0; k < EX_QPT; ++k) {
    const int64_t q = q0 + (int64_t)k * EX_THREADS;
    const bool ok = q < no;
    px[k] = ok ? O[3 * q] : 0.0;
    py[k] = ok ? O[3 * q + 1] : 0.0;
    pz[k] = ok ? O[3 * q + 2] : 0.0;
    acc[k] = 0.0;
  }
  const int64_t fb = (int64_t)blockIdx.y * chunk;
  const int64_t fe = (fb + chunk) < nf ? (fb + chunk) : nf;
  for (int64_t f0 = fb; f0 < fe; f0 += EX_TILE) {
    const int cnt = (int)((fe - f0) < EX_TILE ? (fe - f0) : EX_TILE);
    __syncthreads();
    for (int e = threadIdx.x; e < cnt * 9; e += EX_THREADS) tile[e] = T[9 * f0 + e];
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const double* t = tile + 9 * j;
#pragma unroll
      for (int k = 0; k < EX_QPT; ++k) acc[k] += vos_atan2(t, px[k], py[k], pz[k]);
    }
  }
  for (int k = 0; k < EX_QPT; ++k) {
    const int64_t q = q0 + (int64_t)k * EX_THREADS;
    if (q < no) {
      if (part)
        part[(int64_t)blockIdx.y * no + q] = acc[k];
      else
        W[q] = acc[k] / (2.0 * EX_PI);
    }
  }
}

// W[q] = the chunk partials of query q added in chunk order, divided by 2 pi
// when the terms do not already carry it (3D)
__global__ void k_exact_sum(const double* __restrict__ part, int64_t nchunks, int64_t no, bool div2pi,
                            double* __restrict__ W) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= no) return;
  double s = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) s += part[c * no + q];
  W[q] = div2pi ? s / (2.0 * EX_PI) : s;
}

// Triangle (or edge) chunks per launch: enough blocks for ~16 wavefronts per SIMD
// (256 CUs x 4 SIMDs), chunks of at least 1024 items.
static int64_t exact_chunks(int64_t nf, int64_t qblocks) {
  const int64_t want = (16 * 1024 * 64 / EX_THREADS + qblocks - 1) / qblocks;
  const int64_t most = (nf + 1023) / 1024;
  return want < 1 ? 1 : (want < most ? want : most);
}

// ---- 2D: winding number of a curve of edges (winding_number.cpp:62-89 with
// F.cols() == 2): W(o) = sum over edges (a,b) of signed_angle(a, b, o)
// (signed_angle.cpp:8-41): -atan2(cross, dot) of the unit vectors from a and b
// to o, each term divided by 2 pi.  Terms are summed in edge order within
// each chunk of the edge list; when a small query set splits the edges into
// chunks (exact_chunks), the chunk partials are then added in chunk order, so
// the last bits can differ from the reference's single sum in edge order.
__device__ __forceinline__ double signed_angle_2d(const double* e, double px, double py) {
  double ax = px - e[0], ay = py - e[1], bx = px - e[2], by = py - e[3];
  const double la = sqrt(ax * ax + ay * ay), lb = sqrt(bx * bx + by * by);
  if (la != 0.0) {
    ax /= la;
    ay /= la;
  }
  if (lb != 0.0) {
    bx /= lb;
    by /= lb;
  }
  return -atan2(bx * ay - by * ax, bx * ax + by * ay) / (2.0 * EX_PI);
}

__global__ __launch_bounds__(EX_THREADS) void k_exact2d_gather(const double* __restrict__ V,
                                                                 const int32_t* __restrict__ E, int64_t ne,
                                                                 double* __restrict__ T) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= ne) return;
  for (int k = 0; k < 2; ++k)
    for (int a = 0; a < 2; ++a) T[4 * f + 2 * k + a] = V[2 * (int64_t)E[2 * f + k] + a];
}

// Block (bx, by): queries of block bx against the edges of chunk by, as in
// k_exact; the terms already carry the 1 / 2 pi, so the partials are summed
// as they are.
__global__ __launch_bounds__(EX_THREADS) void k_exact2d(const double* __restrict__ T, int64_t ne, int64_t chunk,
                                                         const double* __restrict__ O, int64_t no,
                                                         double* __restrict__ part, double* __restrict__ W) {
  __shared__ double tile[EX_TILE * 4];
  const int64_t q = (int64_t)blockIdx.x * EX_THREADS + threadIdx.x;
  const bool ok = q < no;
  const double px = ok ? O[2 * q] : 0.0, py = ok ? O[2 * q + 1] : 0.0;
  double acc = 0.0;
  const int64_t fb = (int64_t)blockIdx.y * chunk;
  const int64_t fe = (fb + chunk) < ne ? (fb + chunk) : ne;
  for (int64_t f0 = fb; f0 < fe; f0 += EX_TILE) {
    const int cnt = (int)((fe - f0) < EX_TILE ? (fe - f0) : EX_TILE);
    __syncthreads();
    for (int e = threadIdx.x; e < cnt * 4; e += EX_THREADS) tile[e] = T[4 * f0 + e];
    __syncthreads();
    for (int j = 0; j < cnt; ++j) acc += signed_angle_2d(tile + 4 * j, px, py);
  }
  if (ok) {
    if (part)
      part[(int64_t)blockIdx.y * no + q] = acc;
    else
      W[q] = acc;
  }
}

void exact2d_device(const double* V, const int32_t* E, int64_t ne, const double* O, int64_t no, double* W,
                    hipStream_t st) {
  if (no <= 0) return;
  if (ne <= 0) {
    fill_device<double>(W, no, 0.0, st);
    return;
  }
  const int64_t qblocks = div_up(no, EX_THREADS);
  const int64_t nch = exact_chunks(ne, qblocks);
  GWN_CHECK(nch <= 65535, GWN_ERR_INVALID, "exact 2D: too many edge chunks");
  const int64_t chunk = (ne + nch - 1) / nch;
  // scratch of this call: the edges (4 doubles each), then the partial sums
  const size_t npart = nch > 1 ? (size_t)nch * (size_t)no : 0;
  Scratch ws(scratch_bytes({8 * 4 * (size_t)ne, 8 * npart}), st);
  double* edges = ws.take_n<double>(4 * (size_t)ne);
  double* part = nch > 1 ? ws.take_n<double>(npart) : nullptr;
  hipLaunchKernelGGL(k_exact2d_gather, dim3(div_up(ne, EX_THREADS)), dim3(EX_THREADS), 0, st, V, E, ne, edges);
  check_launch("k_exact2d_gather");
  ProfileScope prof("exact2d", st);
  hipLaunchKernelGGL(k_exact2d, dim3((unsigned)qblocks, (unsigned)nch), dim3(EX_THREADS), 0, st, edges, ne, chunk,
                     O, no, part, W);
  check_launch("k_exact2d");
  if (part) {
    hipLaunchKernelGGL(k_exact_sum, dim3(div_up(no, 256)), dim3(256), 0, st, part, nch, no, false, W);
    check_launch("k_exact_sum");
  }
}

void exact_device(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O, int64_t no, double* W,
                  hipStream_t st) {
  (void)nv;
  if (no <= 0) return;
  if (nf <= 0) {
    fill_device<double>(W, no, 0.0, st);
    return;
  }
  const int64_t qblocks = div_up(no, EX_THREADS * EX_QPT);
  const int64_t nch = exact_chunks(nf, qblocks);
  GWN_CHECK(nch <= 65535, GWN_ERR_INVALID, "exact: too many triangle chunks");
  const int64_t chunk = (nf + nch - 1) / nch;
  // scratch of this call: the triangles (9 doubles each), then the partial sums
  const size_t npart = nch > 1 ? (size_t)nch * (size_t)no : 0;
  Scratch ws(scratch_bytes({8 * 9 * (size_t)nf, 8 * npart}), st);
  double* tris = ws.take_n<double>(9 * (size_t)nf);
  double* part = nch > 1 ? ws.take_n<double>(npart) : nullptr;
  hipLaunchKernelGGL(k_exact_gather, dim3(div_up(nf, EX_THREADS)), dim3(EX_THREADS), 0, st, V, F, nf, tris);
  check_launch("k_exact_gather");
  ProfileScope prof("exact", st);
  hipLaunchKernelGGL(k_exact, dim3((unsigned)qblocks, (unsigned)nch), dim3(EX_THREADS), 0, st, tris, nf, chunk,
                     O, no, part, W);
  check_launch("k_exact");
  if (part) {
    hipLaunchKernelGGL(k_exact_sum, dim3(div_up(no, 256)), dim3(256), 0, st, part, nch, no, true, W);
    check_launch("k_exact_sum");
  }
}


// ---------------------------------------------------------------------------
// The reference's hierarchy (wntree.cpp), one wavefront per query (queries in
// Hilbert order when perm is given).  The wavefront walks the query's
// pre-order records: inside an internal node's closed box it steps to the
// first child; otherwise it sums the node's triangle list (inside a leaf: its
// faces; outside the box: its cap, or its faces when the cap is not smaller)
// and jumps past the subtree.  The 64 lanes share each list, lane l taking
// its triangles l, l+64, ...; each term is solid_angle(...) (solid_angle.cpp:
// atan2 over 2 pi), and the lanes' FP64 partial sums are added in a fixed
// order (wave_sum).  The reference adds a node's children's sums; the orders
// agree to rounding.  A query's result does not depend on the batch it is in.
// One query per lane, wave-coherent or not, ran 6x (100k queries) to 100x
// (1,000 queries) longer: the lists run to dozens of triangles, which a lane
// summed alone while the wavefront's other lanes waited.
__global__ __launch_bounds__(256) void k_wn_tree_lanes(const WNRecord* __restrict__ rec, uint32_t nrec,
                                                        const double* __restrict__ tri,
                                                        const double* __restrict__ O,
                                                        const uint32_t* __restrict__ perm, int64_t no,
                                                        double* __restrict__ W) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)xcd_block() * 4 + (threadIdx.x >> 6);
  if (w >= no) return;
  const int64_t q = perm ? (int64_t)perm[w] : w;
  const double px = O[3 * q], py = O[3 * q + 1], pz = O[3 * q + 2];
  double part = 0.0;
  uint32_t i = 0;
  while (i < nrec) {
    const WNRecord& r = rec[i];
    const bool inside = !(px < r.lo[0] || px > r.hi[0] || py < r.lo[1] || py > r.hi[1] || pz < r.lo[2] ||
                          pz > r.hi[2]);
    if (inside && !r.leaf) {
      ++i;
      continue;
    }
    const uint32_t start = inside ? r.in_start : r.out_start;
    const uint32_t count = inside ? r.in_count : r.out_count;
    for (uint32_t k = (uint32_t)lane; k < count; k += 64)
      part += vos_atan2(tri + 9 * (size_t)(start + k), px, py, pz) / (2.0 * EX_PI);
    i = r.link;
  }
  const double sum = wave_sum(part);
  if (lane == 0) W[q] = sum;
}

__global__ void k_wn_narrow(const double* __restrict__ O, int64_t n3, float* __restrict__ Of) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) Of[i] = (float)O[i];
}

void wntree_upload(const WNTreeHost& T, DevBuf<uint8_t>& rec, DevBuf<double>& tri, hipStream_t st) {
  rec.alloc(sizeof(WNRecord) * T.rec.size());
  if (!T.rec.empty())
    GWN_HIP(hipMemcpyAsync(rec.p, T.rec.data(), sizeof(WNRecord) * T.rec.size(), hipMemcpyHostToDevice, st));
  tri.alloc(T.tri.size());
  if (!T.tri.empty())
    GWN_HIP(hipMemcpyAsync(tri.p, T.tri.data(), sizeof(double) * T.tri.size(), hipMemcpyHostToDevice, st));
  GWN_HIP(hipStreamSynchronize(st));
}

// O no x 3 and W no, device pointers; bbox: the root's box (query order)
void wntree_query_device(const uint8_t* rec, int64_t nrec, const double* tri, const float* bbox, const double* O,
                         int64_t no, double* W, hipStream_t st) {
  if (no <= 0) return;
  if (nrec == 0) {
    fill_device<double>(W, no, 0.0, st);
    return;
  }
  uint32_t* perm = nullptr;
  // small batches walk in input order; larger ones along the Hilbert curve
  // (neighbouring wavefronts then read the same records and triangles)
  const bool sorted = no >= 4096;
  const size_t tb = sorted ? radix_sort_temp_bytes(no, false) : 0;
  Scratch ws(sorted ? scratch_bytes({12 * (size_t)no, 4 * (size_t)no, 4 * (size_t)no, tb}) : 1, st);
  if (sorted) {
    float* Of = ws.take_n<float>(3 * (size_t)no);
    uint32_t* keys = ws.take_n<uint32_t>((size_t)no);
    perm = ws.take_n<uint32_t>((size_t)no);
    void* tmp = ws.take(tb);
    hipLaunchKernelGGL(k_wn_narrow, dim3(div_up(3 * no, 256)), dim3(256), 0, st, O, 3 * no, Of);
    check_launch("k_wn_narrow");
    hilbert_keys(Of, no, bbox, keys, perm, st);
    radix_sort_u32(keys, perm, no, 30, tmp, tb, st);
  }
  ProfileScope prof("exact_tree", st);
  hipLaunchKernelGGL(k_wn_tree_lanes, dim3(div_up(no, 4)), dim3(256), 0, st, reinterpret_cast<const WNRecord*>(rec),
                     (uint32_t)nrec, tri, O, perm, no, W);
  check_launch("k_wn_tree_lanes");
}

}  // namespace gwn
