/*
 * gwn.h -- C ABI of the MI355X generalized-winding-number library (libgwn.so).
 *
 * Drop-in boundary for libigl's winding-number query path.  Each entry point
 * names the reference interface it replaces.  Conventions:
 *   - plain pointers and sizes, row-major contiguous buffers (n x 3);
 *   - caller-owned inputs and outputs; a handle owns its device tree;
 *   - every function returns a gwn_status; gwn_last_error() describes the last
 *     failure on the calling thread (the reference asserts instead);
 *   - *_device entry points take device pointers and a hipStream_t (passed as
 *     void*, NULL = the default stream) and are asynchronous;
 *     the others take host pointers and return when the result is in place;
 *   - queries take a handle by const pointer and are reentrant and thread-safe,
 *     as the reference's const queries are (fast_winding_number.h:191-206):
 *     any number of threads and streams may query one handle at once.  Each
 *     call borrows its temporary device memory from a pool that orders reuse
 *     by stream events, so concurrent calls share no scratch and, once the
 *     pool is warm, a call allocates nothing;
 *   - an empty mesh gives W = 0 (the reference returns an uninitialised value,
 *     FastWindingNumberForSoups.h:7281-7283).
 */
#ifndef GWN_H
#define GWN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GWN_OK = 0,
  GWN_ERR_INVALID = 1,  /* bad argument (null pointer, size, order, index) */
  GWN_ERR_HIP = 2,      /* a HIP runtime call failed */
  GWN_ERR_NOMEM = 3,    /* device allocation failed / over the admission budget */
  GWN_ERR_NODEVICE = 4  /* no usable gfx950 device */
} gwn_status;

/* Tree construction for the triangle soup. */
typedef enum {
  /* The reference topology: UT::BVH<4>::init<BOX_AREA> restated on the host
   * (FastWindingNumberForSoups.h:4579-5944); moments and traversal on the GPU.
   * Results track the reference fast_winding_number to float rounding. */
  GWN_TREE_SAH4 = 0,
  /* Built on the GPU: Morton codes, LDS radix sort, Karras binary radix tree
   * collapsed to a 4-wide tree. */
  GWN_TREE_LBVH = 1
} gwn_tree_kind;

/* Arithmetic of a soup query. */
typedef enum {
  /* FP32 throughout, as the reference (same near/far decisions, same FP32
   * moments); the far field is the reference's Taylor polynomial with its
   * coefficients pre-combined, 1/|q| from the hardware reciprocal square
   * root, and the exact angle uses a fitted polynomial atan2 (about 1e-7
   * rad, the size of atan2f's own rounding) with a hardware reciprocal. */
  GWN_PREC_FP32 = 0,
  /* FP32 far field and FP32 running sum; the VOS solid angles of triangles
   * that may lie within near_dist of the query are computed and summed in
   * FP64 (added to the FP32 sum at the end, in FP64); queries with
   * |w-0.5| < ambiguous_band are then re-evaluated entirely in FP64
   * (moments, far field, atan2, sum). */
  GWN_PREC_MIXED = 1,
  GWN_PREC_FP64 = 2   /* everything in FP64 (moments, far field, atan2) */
} gwn_precision;

typedef struct gwn_soup gwn_soup;   /* igl::FastWindingNumberBVH (fast_winding_number.h:140-152) */
typedef struct gwn_cloud gwn_cloud; /* octree + CM/R/EC of the point-cloud path */
typedef struct gwn_antipodal gwn_antipodal; /* igl::WindingNumberAntipodalScene */

typedef struct {
  int64_t num_vertices, num_triangles;
  int64_t num_nodes;   /* 4-wide nodes */
  int64_t num_slots;   /* non-empty child slots (rope layout) */
  int32_t order, tree_kind, depth;
  int32_t reserved;
  int64_t device_bytes;
  float bbox[6];       /* min xyz, max xyz of the vertices */
} gwn_soup_info;

typedef struct {
  float near_dist;       /* FP64 solid angle for triangles closer than this (default 1e-3) */
  float ambiguous_band;  /* re-evaluate in FP64 when |w-0.5| < band (default 0.05) */
} gwn_mixed_params;

const char* gwn_last_error(void);
int gwn_device_count(void);
/* Library version string and the device it runs on (for logs). */
const char* gwn_version(void);

/* Kernel timing for measurement tools (no counterpart in the reference).
 * While enabled, the library brackets its named kernels ("soup_traverse",
 * "soup_traverse64", "soup_reeval64", "soup_walk_ring" (smaller batches, see gwn_soup_query), "exact",
 * "exact_tree", "cloud_traverse", "radix_sort", "block_order",
 * "closest_point", "knn", "point_areas", "cloud_brute", "exact2d",
 * "curve_traverse") with HIP events
 * recorded on the stream they are launched on; gwn_profile_read synchronises
 * on those events and returns the accumulated time and launch count. */
int gwn_profile_enable(int on);
int gwn_profile_read(const char* kernel, double* total_ms, int64_t* launches);
void gwn_profile_reset(void);
/* One device-to-device copy of `bytes` (a positive multiple of 16) from src
 * to dst by a plain copy kernel (k_copy_calibration), asynchronous on
 * `stream`: a kernel of known traffic (bytes read, bytes written) against
 * which hardware memory counters such as FETCH_SIZE and WRITE_SIZE are
 * scaled. */
int gwn_profile_copy(void* dst, const void* src, int64_t bytes, void* stream);

/* ------------------------------------------------------------------------
 * Exact FP64 generalized winding number.
 * gwn_exact_winding_number replaces igl::winding_number(V,F,O,W),
 * winding_number.h:41-48 (winding_number.cpp:22-56, 3D branch) and follows
 * it: the reference's hierarchy (igl::WindingNumberAABB, built for the call:
 * duplicate vertices merged, median splits down to 100 faces, every node's
 * open boundary closed by a triangle fan whose apex rand() picks, so the
 * call consumes rand() values as the reference's does), then, per origin,
 * solid_angle.cpp:10-55 summed over the faces of the leaves whose boxes
 * contain it and over the fans of the other nodes it meets.  Equal to the sum
 * over every face up to rounding.  V nv x 3, F nf x 3, O no x 3, W no; host
 * arrays.
 * gwn_exact_winding_number_device takes device arrays and sums over every
 * face (the brute-force form of the same quantity; no hierarchy, no rand()).
 */
int gwn_exact_winding_number(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                             int64_t no, double* W);
int gwn_exact_winding_number_device(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                                    int64_t no, double* W, void* stream);
/* Replaces igl::winding_number(V,F,p), winding_number.h:49-60
 * (winding_number.cpp:62-89): solid_angle summed over every face, for each
 * of the no points; host arrays (the same sum as the _device form). */
int gwn_exact_winding_number_bruteforce(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                                        int64_t no, double* W);

/* The hierarchy of gwn_exact_winding_number, built once and queried many
 * times (igl::WindingNumberAABB<double,int> hier(V,F); hier.grow(); then
 * hier.winding_number(p) per origin, WindingNumberAABB.h / WindingNumberTree.h).
 * The build draws rand() as the reference's constructors do.
 * gwn_exact_tree_nodes writes 5 int64 per node in the reference's
 * construction (depth-first pre-) order: #faces, #cap triangles, the fan's
 * apex (an index into the vertices de-duplicated and sorted
 * lexicographically; -1 when the node has no open boundary), left child,
 * right child (-1 for a leaf).  Queries: O no x 3, W no (FP64). */
typedef struct gwn_exact_tree gwn_exact_tree;
int gwn_exact_tree_build(const double* V, int64_t nv, const int32_t* F, int64_t nf, gwn_exact_tree** out);
void gwn_exact_tree_free(gwn_exact_tree* t);
int gwn_exact_tree_info(const gwn_exact_tree* t, int64_t* num_nodes, int64_t* num_triangles);
int gwn_exact_tree_nodes(const gwn_exact_tree* t, int64_t* out);
int gwn_exact_tree_query(const gwn_exact_tree* t, const double* O, int64_t no, double* W);
int gwn_exact_tree_query_device(const gwn_exact_tree* t, const double* O, int64_t no, double* W, void* stream);
/* Host-only (no GPU needed): the same hierarchy built on the host (drawing
 * rand() as gwn_exact_tree_build does); writes up to `capacity` nodes in the
 * gwn_exact_tree_nodes layout and the node count. */
int gwn_exact_tree_host_nodes(const double* V, int64_t nv, const int32_t* F, int64_t nf, int64_t* nodes,
                              int64_t capacity, int64_t* num_nodes);

/* ------------------------------------------------------------------------
 * Triangle soup (UT_SolidAngle).
 * gwn_soup_build replaces igl::fast_winding_number(V,F,order,fwn_bvh),
 * fast_winding_number.h:176-180 (fast_winding_number.cpp:380-416): V is
 * narrowed to float as the reference narrows it; order in {0,1,2}.
 */
int gwn_soup_build(const double* V, int64_t nv, const int32_t* F, int64_t nf, int order, int tree_kind,
                   gwn_soup** out);
int gwn_soup_build_f32(const float* V, int64_t nv, const int32_t* F, int64_t nf, int order, int tree_kind,
                       gwn_soup** out);
void gwn_soup_free(gwn_soup* s);
int gwn_soup_get_info(const gwn_soup* s, gwn_soup_info* info);

/* Replaces igl::fast_winding_number(fwn_bvh, accuracy_scale, Q, W),
 * fast_winding_number.h:191-196 (fast_winding_number.cpp:418-437): each query
 * is narrowed to float (on the device), W = Omega / (4 pi) in double.  The
 * single-point overload (fast_winding_number.h:203-206) is nq = 1.  Batches
 * under 65,536 queries skip the query sort and give each query a wavefront of
 * its own, whose lanes share the query's walk (a query takes about tree-depth
 * many dependent fetches instead of path-length many), so a single point
 * costs one kernel and two small copies.  Larger batches are sorted along a
 * Hilbert curve and walked wave-coherently.  The two walks visit the same
 * records and sum the same terms in different orders: results agree to FP32
 * rounding (FP64 rounding in the FP64 form), and each is deterministic,
 * depending only on the query, the tree, the precision and which side of
 * 65,536 the batch size falls. */
int gwn_soup_query(const gwn_soup* s, const double* Q, int64_t nq, float accuracy_scale, int precision,
                   double* W);
/* Device form: Q float nq x 3, W double nq. */
int gwn_soup_query_device(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                          double* W, void* stream);
int gwn_soup_set_mixed_params(gwn_soup* s, const gwn_mixed_params* p);

/* Scratch-free form of gwn_soup_query_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and bitwise the same results as gwn_soup_query_device; the
 * caller passes device inputs and outputs, and the call's temporary memory is
 * a static array of 22 MiB that the library holds on each device it runs on
 * until the process ends.  It calls no hipMalloc, hipFree or allocating
 * library routine, creates no timing events, and never waits on the host
 * (the mixed form's FP64 re-evaluation is sized on the device).  At most
 * gwn_soup_noalloc_max_queries() = 1,048,576 queries per call
 * (GWN_ERR_INVALID beyond).  Every call on one device shares that device's
 * array: calls that overlap on one device race, so serialise them (one
 * stream, or events between streams). */
int gwn_soup_query_device_noalloc(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale,
                                  int precision, double* W, void* stream);
int64_t gwn_soup_noalloc_max_queries(void);

/* Host-only (no GPU needed): the reference-topology tree of a mesh, i.e.
 * UT_SolidAngle::init's triangle boxes (FastWindingNumberForSoups.h:6566-6590)
 * followed by UT::BVH<4>::init<BOX_AREA> (4579-4634).  Writes up to
 * `capacity` nodes (4 uint32 each, reference encoding) and the node count. */
int gwn_sah4_tree(const float* V, int64_t nv, const int32_t* F, int64_t nf, uint32_t* nodes, int64_t capacity,
                  int64_t* num_nodes);

/* Introspection for parity checks: the 4-wide node array in the reference
 * encoding (uint32 child[4]: item index, 0x80000000|node, 0xFFFFFFFF empty,
 * FastWindingNumberForSoups.h:4217-4232) and the per-node BoxData, 92 floats in
 * the reference member order (6477-6521). */
int gwn_soup_copy_nodes(const gwn_soup* s, uint32_t* out);
int gwn_soup_copy_boxdata(const gwn_soup* s, float* out);

/* ------------------------------------------------------------------------
 * Point cloud.
 * gwn_cloud_build replaces igl::octree(P,point_indices,CH,CN,W) (octree.h,
 * octree.cpp:7-169) followed by the precompute
 * igl::fast_winding_number(P,N,A,point_indices,CH,order,CM,R,EC),
 * fast_winding_number.h:44-53.  P, N n x 3, A n; FP32.  With N and A both
 * NULL it builds the octree alone (igl::octree): such a cloud has no
 * expansion and cannot be queried.
 */
int gwn_cloud_build(const float* P, const float* N, const float* A, int64_t n, int order, gwn_cloud** out);
/* The precompute (fast_winding_number.h:44-53, .cpp:8-124) and the
 * evaluation inputs (fast_winding_number.h:88-99) over a caller's octree:
 * CH ncells x 8 (int32, -1 for leaves), and point_indices flattened as
 * cell_start (ncells + 1 entries, cell_start[0] = 0) and point_list (the
 * lists back to back, in the caller's order).  CM, R, EC (ncells x 3, ncells,
 * ncells x (3 | 12 | 39)) are used as given when all three are non-NULL and
 * computed as the reference's precompute computes them when all are NULL
 * (centres summed in FP32 over each cell's list in list order).  The cells
 * reachable from cell 0 must form a tree whose split cells have 8 children
 * (GWN_ERR_INVALID otherwise).  gwn_cloud_copy_octree is refused on such a
 * cloud (the caller holds the octree); gwn_cloud_copy_expansion returns CM,
 * R, EC. */
int gwn_cloud_build_from_octree(const float* P, const float* N, const float* A, int64_t n, const int32_t* CH,
                                int64_t ncells, const int64_t* cell_start, const int32_t* point_list, int order,
                                const float* CM, const float* R, const float* EC, gwn_cloud** out);
void gwn_cloud_free(gwn_cloud* c);
/* number of octree cells (CH rows) */
int gwn_cloud_num_cells(const gwn_cloud* c, int64_t* ncells);
/* CH ncells x 8 (int32, -1 for leaves), CN ncells x 3, W ncells, and the
 * points of each cell (point_indices in the reference): cell c holds
 * point_list[cell_start[c] .. cell_start[c] + cell_count[c]), point_list being
 * the n point indices in octree order (sort a cell's slice to get the
 * reference's ascending point_indices list).
 * Any output pointer may be NULL. */
int gwn_cloud_copy_octree(const gwn_cloud* c, int32_t* CH, float* CN, float* W, int64_t* cell_start,
                          int32_t* cell_count, int32_t* point_list);
int gwn_cloud_copy_expansion(const gwn_cloud* c, float* CM, float* R, float* EC);
/* Replaces igl::fast_winding_number(P,N,A,point_indices,CH,CM,R,EC,Q,beta,WN),
 * fast_winding_number.h:88-99; beta <= 0 is the reference's brute-force
 * branch (fast_winding_number.cpp:300-310).  Q nq x 3 float, WN float. */
int gwn_cloud_query(const gwn_cloud* c, const float* Q, int64_t nq, double beta, float* WN);
int gwn_cloud_query_device(const gwn_cloud* c, const float* Q, int64_t nq, double beta, float* WN, void* stream);

/* ------------------------------------------------------------------------
 * Antipodal method (WindingNumberAntipodalScene.h, winding_number_antipodal.h:
 * 17-53): W(p) = sum over the weighted open-boundary edges of the mesh of
 * half the solid angle of (-x0, a - p, b - p) / (2 pi), plus the signed
 * number of triangles the ray p + t x0 (t >= 0) crosses (+1 through a
 * triangle's front, (v1-v0)x(v2-v0) . x0 > 0).  The reference counts the
 * crossings with a caller-supplied intersector (Embree,
 * embree/EmbreeIntersector.cpp:390-443); here they are counted on the GPU
 * over the mesh's own tree, in FP64, with a watertight ray-triangle test.
 * gwn_antipodal_build is the scene constructor (x0 NULL: the reference's
 * default normalize(1, sqrt 2, sqrt 3)); gwn_antipodal_query its batch
 * winding_number; gwn_winding_number_antipodal the one-shot
 * igl::winding_number_antipodal(V,F,intersector,O,W).  V nv x 3, F nf x 3,
 * O no x 3, W no; FP64. */
int gwn_antipodal_build(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* x0,
                        gwn_antipodal** out);
void gwn_antipodal_free(gwn_antipodal* a);
int gwn_antipodal_info(const gwn_antipodal* a, int64_t* num_boundary_segments, int64_t* num_faces, double* x0);
int gwn_antipodal_query(const gwn_antipodal* a, const double* O, int64_t no, double* W);
int gwn_antipodal_query_device(const gwn_antipodal* a, const double* O, int64_t no, double* W, void* stream);
int gwn_winding_number_antipodal(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                                 int64_t no, double* W);

/* ------------------------------------------------------------------------
 * Point areas for a cloud without them: k nearest neighbours and a local
 * Delaunay star (the pre-step of copyleft/cgal/fast_winding_number.cpp:32-43).
 *
 * gwn_knn replaces igl::knn(P,V,k,point_indices,CH,CN,W,I), knn.h:51-67
 * (knn.cpp:24-127); igl::knn(P,k,point_indices,CH,CN,W,I) (knn.h:33-45) is
 * V = P.  Row i of I lists the k candidates nearest to P.row(i), nearest
 * first (FP64 distances; equal distances in ascending index order).  The
 * candidates are cand[0 .. ncand) (indices into V): the first point of every
 * non-empty leaf of the octree, the only points the reference's search
 * reaches; cand = NULL makes every point of V a candidate.  When nv <= k the
 * reference returns all of V sorted (knn.cpp:46-55) and so does this: I has
 * gwn_knn_columns(nv, k) = min(nv, k) columns, row-major int32.
 * GWN_ERR_INVALID when there are fewer than k candidates (where the
 * reference's queue runs dry) or a coordinate is not finite. */
int gwn_knn_columns(int64_t nv, int64_t k);
int gwn_knn(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand, int64_t ncand,
            int32_t* I);
int gwn_knn_device(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand,
                   int64_t ncand, int32_t* I, void* stream);
/* gwn_point_areas replaces igl::copyleft::cgal::point_areas(P,I,N,A,T),
 * point_areas.h:44-62 (point_areas.cpp:37-170): per point, the neighbours
 * I.row(i) whose normal has a positive dot product with the first
 * neighbour's, a least-squares plane through them (T = its unit normal, on
 * the side of N.row(i)), and the area the reference sums over the Delaunay
 * triangles of the projected neighbours at the first kept neighbour.  Fewer
 * than 3 kept neighbours, or a non-finite sum, give A = 0 and (then) T = 0.
 * The reference triangulates with CGAL; this builds the same star directly.
 * N may be NULL (no normal filter, T unoriented); T may be NULL.  P, N n x 3,
 * I n x k (indices into P), A n, T n x 3. */
int gwn_point_areas(const double* P, int64_t n, const int32_t* I, int64_t k, const double* N, double* A,
                    double* T);
int gwn_point_areas_device(const double* P, int64_t n, const int32_t* I, int64_t k, const double* N, double* A,
                           double* T, void* stream);
/* gwn_cloud_build_estimated_areas is the build half of
 * igl::copyleft::cgal::fast_winding_number(P,N,Q,expansion_order,beta,WN)
 * (copyleft/cgal/fast_winding_number.h, fast_winding_number.cpp:16-43): the
 * areas come from knn with k = 21 over the candidates of P's own octree (all
 * points but exact copies of a lower-indexed one) and point areas, then the
 * cloud is built as gwn_cloud_build builds it (FP32).  A (n doubles, may be
 * NULL) receives the estimated areas.  Query it with gwn_cloud_query. */
int gwn_cloud_build_estimated_areas(const double* P, const double* N, int64_t n, int order, double* A,
                                    gwn_cloud** out);

/* ------------------------------------------------------------------------
 * Signed distance (signed_distance.h:61-89, signed_distance.cpp:28-140):
 * closest point on the soup (FP64, from the vertices as given) and a sign
 * from the winding number.  Sign types carry the reference's enum values.
 */
typedef enum {
  GWN_SIGN_PSEUDONORMAL = 0,        /* s = sign((q - c) . n), n the angle-weighted pseudonormal at c */
  GWN_SIGN_WINDING_NUMBER = 1,      /* s = 1 - 2 w, w the exact winding number */
  GWN_SIGN_DEFAULT = 2,             /* as WINDING_NUMBER */
  GWN_SIGN_UNSIGNED = 3,            /* s = 1 */
  GWN_SIGN_FAST_WINDING_NUMBER = 4  /* s = 1 - 2 |w|, w = fast winding number at accuracy_scale 2 */
} gwn_sign_type;

/* igl::signed_distance(P,V,F,sign_type,lower_bound,upper_bound,S,I,C,N) over
 * the soup's mesh (signed_distance.cpp:28-140; the FWN sign uses the soup's
 * tree and expansion order; the reference builds an order-2 tree).  P np x 3;
 * S np (s * distance), I np (closest face), C np x 3 (closest point).  A point
 * whose squared distance lies outside [low^2, up^2) derived from the bounds as
 * signed_distance.cpp:79-84 gets S = NaN, I = nf + 1, C = 0; the reference's
 * unbounded overload (signed_distance.h:81-89) is lower = DBL_MIN,
 * upper = DBL_MAX.  An empty mesh gives S = NaN, I = -1, C = NaN.
 * WINDING_NUMBER / DEFAULT take w through the exact hierarchy built for the
 * call, as the reference's does (and synchronise the stream).  PSEUDONORMAL
 * builds the face, angle-weighted vertex and edge normals of the mesh on the
 * first such call (pseudonormal_test.cpp:24-141 picks among them); N np x 3,
 * when not NULL, receives the normal used (0 for points out of bounds, whose
 * rows the reference leaves unset); for other sign types N is not written. */
int gwn_signed_distance_with_normals(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower_bound,
                                     double upper_bound, double* S, int32_t* I, double* C, double* N);
int gwn_signed_distance_with_normals_device(gwn_soup* s, const double* P, int64_t np, int sign_type,
                                            double lower_bound, double upper_bound, double* S, int32_t* I, double* C,
                                            double* N, void* stream);
/* The same without N. */
int gwn_signed_distance(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower_bound,
                        double upper_bound, double* S, int32_t* I, double* C);
int gwn_signed_distance_device(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower_bound,
                               double upper_bound, double* S, int32_t* I, double* C, void* stream);

/* 2D curves: igl::winding_number(V,F,O,W) with F.cols() == 2
 * (winding_number.cpp:22-33, 62-89; signed_angle.cpp:8-41), FP64 on the GPU.
 * V nv x 2, E ne x 2 (edges), O no x 2, W no; host buffers. */
int gwn_exact_winding_number_2d(const double* V, int64_t nv, const int32_t* E, int64_t ne, const double* O,
                                int64_t no, double* W);

/* ------------------------------------------------------------------------
 * 2D fast winding number of a curve: UT_SubtendedAngle<float,float>
 * (FastWindingNumberForSoups.h:6315-6370).  gwn_curve_build replaces
 * UT_SubtendedAngle::init(nsegments, segment_points, npoints, positions,
 * order) (7342-7604): P np x 2 float, E ne x 2 int32 (the segment point
 * indices), order in {0,1,2}; tree_kind as for the soup (GWN_TREE_SAH4 is the
 * reference topology, BVH<4>::init<BOX_AREA,S,2>).  The handle keeps its own
 * device copies (the reference keeps the caller's pointers).
 * gwn_curve_query replaces computeAngle(query_point, accuracy_scale)
 * (7700-7828) for nq queries (Q nq x 2 float): angle[i] is the signed angle
 * the curve subtends at Q[i] (the winding number is angle / 2 pi); an empty
 * curve gives 0.
 */
typedef struct gwn_curve gwn_curve;
typedef struct {
  int64_t num_points, num_segments;
  int64_t num_nodes, num_slots;
  int32_t order, tree_kind, depth, reserved;
  int64_t device_bytes;
  float bbox[4]; /* min xy, max xy of the points */
} gwn_curve_info;

int gwn_curve_build(const float* P, int64_t np, const int32_t* E, int64_t ne, int order, int tree_kind,
                    gwn_curve** out);
void gwn_curve_free(gwn_curve* c);
int gwn_curve_get_info(const gwn_curve* c, gwn_curve_info* info);
int gwn_curve_query(const gwn_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle);
/* Device form: Q and angle are device pointers. */
int gwn_curve_query_device(const gwn_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle,
                           void* stream);
/* Introspection for parity checks: the node array (reference encoding, as
 * gwn_soup_copy_nodes) and the per-node BoxData, 48 floats in the reference
 * member order (7287-7320). */
int gwn_curve_copy_nodes(const gwn_curve* c, uint32_t* out);
int gwn_curve_copy_boxdata(const gwn_curve* c, float* out);
/* Host-only (no GPU needed): the reference-topology tree of a curve, i.e. the
 * segment boxes of UT_SubtendedAngle::init (7365-7390) followed by
 * BVH<4>::init<BOX_AREA,S,2> (7392).  As gwn_sah4_tree. */
int gwn_sah4_tree_2d(const float* P, int64_t np, const int32_t* E, int64_t ne, uint32_t* nodes, int64_t capacity,
                     int64_t* num_nodes);

/* ------------------------------------------------------------------------
 * Mesh files (read_triangle_mesh.h; readOFF.cpp, readOBJ.cpp,
 * polygons_to_triangles.cpp): an .off or .obj file (chosen by extension) as a
 * triangle mesh, polygons split into fans (p0, pc, pc+1).  *V (nv x 3 double)
 * and *F (nf x 3 int32, 0-based, indices as in the file) are allocated here;
 * release them with gwn_free.
 */
int gwn_read_triangle_mesh(const char* path, double** V, int64_t* nv, int32_t** F, int64_t* nf);
void gwn_free(void* p);

/* ------------------------------------------------------------------------
 * Serving: continuous batching of soup queries against resident meshes (no
 * counterpart in the reference, whose callers such as
 * signed_distance.cpp:50-121 query one point at a time on the host).
 *
 * A server owns `slots` HIP streams, each served by its own host thread, and a
 * device-memory budget.  Meshes are built once and stay resident (their
 * device bytes count against the budget; add_mesh fails with GWN_ERR_NOMEM
 * when they do not fit).  gwn_server_submit returns at once; a request is
 * admitted, first come first served, when a slot is idle and its device
 * footprint (gwn_server_request_bytes: queries, results and sort workspace of
 * one chunk of at most 1,048,576 queries, doubled for larger requests, which run
 * as a pipeline of such chunks) fits in the remaining budget.
 * Admitted requests run concurrently on different streams, so their kernels
 * overlap on the device; HIP maps a process's streams onto GPU_MAX_HW_QUEUES
 * hardware queues (4 unless set before the process first uses HIP), and slots
 * that share a queue run one after the other, so set it to at least `slots`.
 * A request's result is written to the caller's W
 * (which must stay valid until the request is done) and is bit-identical to
 * gwn_soup_query on the same mesh and queries.  Timestamps are seconds on a
 * steady clock shared by all requests of a server.
 */
typedef struct gwn_server gwn_server;
typedef struct {
  double t_submit, t_start, t_done; /* arrival, admission, result in host memory */
  int64_t num_queries;
  int32_t mesh, status;             /* status: GWN_OK or the request's error */
  double t_uploaded;                /* last chunk narrowed and its copy to the device enqueued */
  double t_device;                  /* last chunk's results back in pinned host memory */
} gwn_request_info;

int gwn_server_create(int slots, int64_t device_budget_bytes, gwn_server** out);
void gwn_server_destroy(gwn_server* s);
/* builds the mesh's tree (and FP64 moments) and keeps it resident */
int gwn_server_add_mesh(gwn_server* s, const double* V, int64_t nv, const int32_t* F, int64_t nf, int order,
                        int tree_kind, int32_t* mesh_id);
/* the same for a mesh file (gwn_read_triangle_mesh) */
int gwn_server_add_mesh_file(gwn_server* s, const char* path, int order, int tree_kind, int32_t* mesh_id);
int gwn_server_submit(gwn_server* s, int32_t mesh_id, const double* Q, int64_t nq, float accuracy_scale,
                      int precision, double* W, int64_t* request_id);
/* blocks until the request is done; returns its status */
int gwn_server_wait(gwn_server* s, int64_t request_id);
/* blocks until every submitted request is done */
int gwn_server_drain(gwn_server* s);
int gwn_server_request_info(const gwn_server* s, int64_t request_id, gwn_request_info* info);
/* device bytes: resident meshes, slot buffers in use, budget */
int gwn_server_memory(const gwn_server* s, int64_t* resident, int64_t* in_flight, int64_t* budget);
/* seconds on the server clock (for callers that time arrivals) */
double gwn_server_now(const gwn_server* s);
/* device bytes a request of num_queries needs while it runs */
int64_t gwn_server_request_bytes(int64_t num_queries);
/* sizes every slot's device buffer for requests of num_queries now, so that
 * no slot allocates while serving; needs an idle server (no request in
 * flight); GWN_ERR_NOMEM when the buffers do not fit in the budget */
int gwn_server_reserve(gwn_server* s, int64_t num_queries);

#ifdef __cplusplus
}
#endif
#endif
