// gwn_internal.h -- shared definitions of the HIP winding-number library.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gwn.h"

struct gwn_soup;
struct gwn_curve;

namespace gwn {

// ---- error reporting (thread-local last error, returned as gwn_status) ----
void set_error(const std::string& msg);
struct Error {
  int code;
  std::string msg;
};

#define GWN_HIP(call)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      throw ::gwn::Error{GWN_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " (" + \
                                          __FILE__ + ":" + std::to_string(__LINE__) + ")"};             \
  } while (0)

#define GWN_CHECK(cond, code, msg)                           \
  do {                                                       \
    if (!(cond)) throw ::gwn::Error{(code), std::string(msg)}; \
  } while (0)

// ---- device buffer owning hipMalloc'ed memory ----
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Error{GWN_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
    }
    n = count;
  }
  // grow-only (keeps capacity); contents not preserved
  void reserve(size_t count) {
    if (count > n || p == nullptr) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---- per-call device scratch (scratch.cpp) ----
// Borrows an idle pooled buffer of at least `bytes` for work enqueued on
// `st` (the stream is made to wait for the buffer's previous user); the
// destructor returns it, recording an event on `st` that orders the next
// borrower after this call's work.  No host synchronisation; no allocation
// once the pool holds a buffer of the size class.  take() carves 256-byte
// aligned sub-buffers in order.
struct ScratchEntry;
class Scratch {
 public:
  Scratch(size_t bytes, hipStream_t st);
  ~Scratch();
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  uint8_t* take(size_t bytes);
  template <typename T>
  T* take_n(size_t count) {
    return reinterpret_cast<T*>(take(sizeof(T) * count));
  }
  uint8_t* p = nullptr;
  size_t capacity = 0;

 private:
  ScratchEntry* e_ = nullptr;
  hipStream_t st_ = nullptr;
  size_t off_ = 0;
};
// bytes that take() calls of these sizes use together (with the alignment)
inline size_t scratch_bytes(std::initializer_list<size_t> parts) {
  size_t t = 0;
  for (size_t b : parts) t = (t + 255) / 256 * 256 + b;
  return t;
}

inline unsigned div_up(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Checks that a kernel launch was accepted (async errors surface at the sync).
inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error{GWN_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

void ensure_device();

// p[0 .. n) = v.  Used instead of hipMemsetAsync so that a caller's captured
// graph replays the fill.
template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
template <typename T>
inline void fill_device(T* p, int64_t n, T v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<T>, dim3(div_up(n, 256)), dim3(256), 0, st, p, n, v);
  check_launch("k_fill");
}

// Brackets a kernel launch with HIP events when profiling is enabled
// (gwn_profile_enable); the elapsed time is collected by gwn_profile_read.
class ProfileScope {
 public:
  ProfileScope(const char* name, hipStream_t st, bool enabled = true);
  ~ProfileScope();
  ProfileScope(const ProfileScope&) = delete;
  ProfileScope& operator=(const ProfileScope&) = delete;

 private:
  const char* name_;
  hipStream_t st_;
  hipEvent_t a_{}, b_{};
  bool active_ = false;
};

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin by
// blockIdx.x.  xcd_block() renumbers them so that XCD x runs a contiguous range
// of logical blocks: with queries in Morton order each XCD then walks one
// spatial region and its own L2 holds that region's part of the tree.
__device__ __forceinline__ uint32_t xcd_block() {
  constexpr uint32_t kXcd = 8;
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t q = nb / kXcd, r = nb % kXcd, x = b % kXcd, k = b / kXcd;
  return x * q + (x < r ? x : r) + k;
}

// ---- tree encoding shared with the reference (FastWindingNumberForSoups.h:4217-4232) ----
constexpr uint32_t NODE_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t NODE_INTERNAL = 0x80000000u;

// Host build of the reference topology (UT::BVH<4>::init<BOX_AREA>).
// boxes: nf x 6 floats (xmin,xmax,ymin,ymax,zmin,zmax); returns nodes (4 per node).
// naxes = 2 builds the tree of 2D boxes (z extents ignored), as
// UT_SubtendedAngle::init does (FastWindingNumberForSoups.h:7392).
void sah4_build(const float* boxes, uint32_t nboxes, std::vector<uint32_t>& nodes, int naxes = 3);

// ---- radix sort (LDS), device ----
// Stable sort of (keys, values) by key; keys uint32 or uint64 with the low
// `bits` bits significant.  tmp buffers are sized by radix_sort_temp_bytes.
size_t radix_sort_temp_bytes(int64_t n, bool key64);
void radix_sort_u32(uint32_t* keys, uint32_t* vals, int64_t n, int bits, void* tmp, size_t tmp_bytes,
                    hipStream_t st);
void radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int bits, void* tmp, size_t tmp_bytes,
                    hipStream_t st);

// ---- device scans ----
size_t scan_temp_bytes(int64_t n);
// exclusive scan of uint32 -> uint32 (in place allowed), total written to *total_dev if non-null
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, void* tmp,
                        hipStream_t st);

// ---- LBVH build on the device: nodes in the reference encoding, DFS preorder ----
// tri_boxes: nf x 6 floats on device. Writes nodes (device, 4 per node) and returns count.
int64_t lbvh_build(const float* tri_boxes_dev, int64_t nf, const float* bbox_host, DevBuf<uint32_t>& nodes,
                   hipStream_t st);

// Nodes of a 4-wide tree (reference encoding) grouped by depth: order[]
// lists the nodes of depth d at offset[d] .. offset[d+1].
struct Levels {
  std::vector<uint32_t> offset;
  DevBuf<int32_t> order;
  int maxdepth = 0;
};
void compute_levels(const uint32_t* nodes, int64_t nn, Levels& L, hipStream_t st);
// First slot of every node in the rope (depth-first slot) layout, given the
// slot count of each node's subtree (nsub); first[0] = 0.
void rope_first_slots(const Levels& L, const uint32_t* nodes, const uint32_t* nsub, uint32_t* first,
                      hipStream_t st);
// soup queries (soup.hip)
size_t soup_query_work_bytes(int64_t nq);
void soup_ensure_dist(gwn_soup* s, hipStream_t st);
// .off / .obj file as a triangle mesh (polygons split into fans)
void read_triangle_mesh(const char* path, std::vector<double>& V, std::vector<int32_t>& F);
void morton_keys(const float* Q, int64_t nq, const float* bbox, uint32_t* keys, uint32_t* vals, hipStream_t st);
// 30-bit Hilbert keys (10 bits per axis over bbox) + identity values: query order for the tree walks
void hilbert_keys(const float* Q, int64_t nq, const float* bbox, uint32_t* keys, uint32_t* vals, hipStream_t st);
// Costliest-first order of the blocks of 256 sorted queries (perm) within
// each XCD's range of logical blocks, block costs read from a kCostGrid^3
// grid of log-scale walk costs over bbox: order ceil(nq/256) entries and
// cost ceil(nq/256) bytes of workspace.
void block_order(const uint8_t* grid, const float* bbox, const float* Q, const uint32_t* perm, int64_t nq,
                 uint32_t* order, uint8_t* cost, hipStream_t st);
void signed_distance_device(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower, double upper,
                            double* S, int32_t* I, double* C, double* N, hipStream_t st);
// the normals of the pseudonormal sign test, in the reference's summation
// order (normals.cpp)
void pseudonormals_host(const double* V, int64_t nv, const int32_t* F, int64_t nf, std::vector<double>& FN,
                        std::vector<double>& VN, std::vector<double>& ENf);
// k nearest neighbours and point areas (knn.hip)
int knn_columns(int64_t nv, int64_t k);
void knn_device(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand,
                int64_t ncand, int32_t* I, hipStream_t st);
void point_areas_device(const double* P, int64_t n, const int32_t* I, int64_t k, const double* N, double* A,
                        double* T, hipStream_t st);
int64_t distinct_points_device(const double* V, int64_t n, DevBuf<int32_t>& cand, hipStream_t st);
void to_float_device(const double* x, int64_t n, float* y, hipStream_t st);
// 2D curve (curve.hip): device setup + tree + moments, and queries
void curve_build_device(gwn_curve* c, hipStream_t st);
void curve_query_device(const gwn_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle,
                        hipStream_t st);
// traversal kernel: WALK_WAVE (wave-coherent over the Hilbert-sorted batch)
// and WALK_WAVE_HANDOVER (the same until a wavefront's lanes diverge, then
// per-lane: shorter batches on large trees with thin-triangle fans, slower on
// coherent ones) give bitwise the same results (sequential sums in rope
// order).  WALK_RING (one wavefront per query, in input order, no workspace)
// sums per lane and then across lanes: deterministic, and within rounding of
// the sequential walks.
constexpr int WALK_WAVE = 0, WALK_WAVE_HANDOVER = 3, WALK_RING = 4;
constexpr int64_t kRingMaxQueries = 65536;
constexpr int kCostGrid = 32;
// Walk of a batch of nq queries on tree s, used by every entry point and by
// the serving engine (on the whole request), so that a batch's result does not
// depend on the path it took: the ring walk below kRingMaxQueries, the
// wave-coherent walk above; latency_bound callers get its hand-over form on
// trees of 2^17+ slots.  (The ring walk, over Hilbert-sorted queries, also
// finishes batches of up to ~500k queries sooner on large trees when run
// alone, but it keeps the whole GPU busy for it: under the serving load it
// cost throughput and latency, p50 30 ms against 11 ms at scale 4.9.)
int soup_walk_for(const gwn_soup* s, int64_t nq, bool latency_bound);
// flags of soup_query_device_ws: WS_DEVICE_COUNT sizes the mixed form's FP64
// re-evaluation on the device (no host round trip); WS_NO_PROFILE creates no
// kernel-timing events
constexpr int WS_DEVICE_COUNT = 1, WS_NO_PROFILE = 2;
void soup_query_device_ws(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                          double* W, hipStream_t st, uint8_t* workspace, size_t ws_bytes, int walk, int flags = 0);
// scratch-free soup query over the static device scratch (soup.hip)
void soup_query_device_noalloc(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                               double* W, hipStream_t st);
int64_t soup_noalloc_max_queries();

}  // namespace gwn

struct gwn_soup {
  int64_t nv = 0, nf = 0;
  int order = 2, tree_kind = 0, depth = 0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};
  gwn::DevBuf<float> U;         // nv x 3
  gwn::DevBuf<int32_t> F;       // nf x 3
  gwn::DevBuf<float> tri;       // nf x 9 (a, b, c) for exact evaluation
  gwn::DevBuf<float> boxes;     // nf x 6 triangle boxes
  gwn::DevBuf<uint32_t> nodes;  // nnodes x 4
  int64_t nnodes = 0;
  gwn::DevBuf<float> boxdata;   // nnodes x 92, reference BoxData order
  gwn::DevBuf<float4> slots;    // nslots x 6 (rope layout, FP32 moments)
  int64_t nslots = 0;
  gwn::DevBuf<double> slots64;  // nslots x 24 (FP64 moments), built with the tree
  gwn::DevBuf<double> V64;      // nv x 3, the vertices as given (closest-point queries)
  gwn::DevBuf<float4> dbox;     // nslots x 2 slot boxes for closest-point queries, built on first use
  // pseudonormal sign: face normals nf x 3, angle-weighted vertex normals
  // nv x 3, edge normals per face and corner nf x 3 x 3; built on first use
  gwn::DevBuf<double> pn_face, pn_vertex, pn_edge;
  gwn_mixed_params mixed{1e-3f, 0.05f};
  // Walk cost of a query at the centre of each cell of a kCostGrid^3 grid over
  // bbox (records visited at accuracy scale 2, as a log-scale byte), built
  // with the tree: the wave walks start their costliest blocks first.
  gwn::DevBuf<uint8_t> cost_grid;
  // After the build a handle is read-only to queries, except for dbox and the
  // pseudonormals, which the first query that needs them builds under
  // dist_mu / pn_mu (dist_ready / pn_ready set after).
  std::mutex dist_mu, pn_mu;
  std::atomic<bool> dist_ready{false}, pn_ready{false};
};

// 2D curve (UT_SubtendedAngle<float,float>): segments, the 4-wide tree over
// their boxes, the reference BoxData and the rope-ordered slot records.
struct gwn_curve {
  int64_t np = 0, ne = 0;
  int order = 2, tree_kind = 0, depth = 0;
  float bbox[4] = {0, 0, 0, 0};  // min xy, max xy of the points
  gwn::DevBuf<float> P;          // np x 2
  gwn::DevBuf<int32_t> E;        // ne x 2
  gwn::DevBuf<float4> seg;       // ne: a.xy, b.xy
  gwn::DevBuf<float> boxes;      // ne x 6 segment boxes (z = 0)
  gwn::DevBuf<uint32_t> nodes;   // nnodes x 4
  int64_t nnodes = 0;
  gwn::DevBuf<float> boxdata;    // nnodes x 48, reference BoxData order
  gwn::DevBuf<float4> slots;     // nslots x 3 (rope layout)
  int64_t nslots = 0;
};

// Antipodal-method scene: the soup's triangles, tree and slot boxes for the
// rays, and the weighted open-boundary segments.
struct gwn_antipodal {
  gwn_soup soup;
  gwn::DevBuf<double> seg;  // nseg x 8: a.xyz, b.xyz, w, 0
  int64_t nseg = 0;
  double x0[3] = {0, 0, 0};
};
// Exact winding-number hierarchy (igl::WindingNumberAABB): pre-order records
// and triangle corners on the device; the per-node counts kept for parity
// checks.
struct gwn_exact_tree {
  int64_t nv = 0, nf = 0, nrec = 0, ntri = 0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};  // root box (query order)
  gwn::DevBuf<uint8_t> rec;             // nrec WNRecord (wntree.h)
  gwn::DevBuf<double> tri;              // ntri x 9
  std::vector<int64_t> nodes;           // nrec x 5: #faces, #cap, apex, left, right
};
namespace gwn {
// exact.hip: the hierarchy's records on the device, and its queries (O, W device)
void wntree_query_device(const uint8_t* rec, int64_t nrec, const double* tri, const float* bbox, const double* O,
                         int64_t no, double* W, hipStream_t st);
// igl::WindingNumberAABB<double,int> of host arrays V, F (built on the host,
// drawing rand() as the reference does; records uploaded to the device)
void exact_tree_build(gwn_exact_tree* t, const double* V, int64_t nv, const int32_t* F, int64_t nf);
}  // namespace gwn

struct gwn_cloud {
  int64_t n = 0, ncells = 0, nrec = 0;
  int order = 2, nterms = 39;
  bool root_leaf = true;               // the root has no children: brute force over its points
  int64_t root_start = 0, root_count = 0;  // the root leaf's points in Ps/NAs
  bool user_octree = false;            // built over a caller's octree (no CN/W)
  bool octree_only = false;            // built without normals/areas: no expansion, no queries
  float bbox[6] = {0, 0, 0, 0, 0, 0};
  gwn::DevBuf<float> P, N, A;          // n x 3, n x 3, n (original order)
  gwn::DevBuf<float> Ps, NAs;          // points in octree order: xyz and N*A
  gwn::DevBuf<int32_t> perm;           // octree position -> original point index
  gwn::DevBuf<int32_t> CH;             // ncells x 8
  gwn::DevBuf<float> CN, Wd;           // ncells x 3, ncells
  gwn::DevBuf<int64_t> cstart;         // ncells: first point (octree order)
  gwn::DevBuf<int32_t> ccount;         // ncells: number of points
  gwn::DevBuf<float> CM, R, EC;        // ncells x 3, ncells, ncells x nterms
  gwn::DevBuf<float4> head;            // nrec: CM.xyz, R of each non-empty non-root cell (rope order)
  gwn::DevBuf<uint2> meta;             // nrec: (skip, cell) or (LEAF|first point, count)
  gwn::DevBuf<float4> poly;            // nrec x 5: the cell's expansion as one polynomial (internal records)
  // records visited (and leaf points summed) by a query at each cell centre
  // of a kCostGrid^3 grid over bbox, at beta 2, as a log-scale byte: the
  // walk's block order (empty when the root is a leaf)
  gwn::DevBuf<uint8_t> cost_grid;
};
