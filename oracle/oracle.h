/*
 * oracle.h -- CPU restatement of libigl's generalized-winding-number path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in libigl_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function restates the reference algorithm step by step (same tree
 * topology, same operation order, no FMA contraction when built with
 * -ffp-contract=off) and cites the reference file:line it follows.
 *
 * Pinning (what checks each function against something independent of it):
 *   oracle_solid_angle / oracle_winding_number_bf
 *       closed forms: closed mesh inside = 1 / outside = 0, single-tet signed
 *       distance case (tests/include/igl/signed_distance.cpp:4-44), orientation
 *       flip w -> -w, hemisphere cap values; and the compiled reference
 *       UTsignedSolidAngleTri in double (oracle/_ref).
 *   oracle_soup_*
 *       bit-exact against the compiled reference UT_SolidAngle (oracle/_ref):
 *       node arrays, BoxData and per-query Omega; plus closed-mesh invariants.
 *   oracle_curve_*
 *       bit-exact against the compiled reference UT_SubtendedAngle
 *       (oracle/_ref): node arrays, BoxData and per-query angles; plus closed
 *       curves (2 pi inside, 0 outside) and the exact segment angle sum.
 *   oracle_point_tri_sqrdist / oracle_signed_distance
 *       closed-form closest points in every Voronoi region of a right
 *       triangle, a dense barycentric sampling bound, the single-tet golden
 *       value of tests/include/igl/signed_distance.cpp:4-44 (sqrt(1/12)) and
 *       sign checks on closed meshes.
 *   oracle_winding_number_antipodal
 *       equals the exact winding number (oracle_winding_number_bf) to 1e-9 on
 *       closed meshes (where it is an integer, the boundary being empty),
 *       open meshes, a single triangle and random non-manifold soups (the
 *       reference's own test, tests/include/igl/embree/
 *       winding_number_antipodal.cpp, asserts 1e-5 there), for the default
 *       and for axis-aligned directions.
 *   oracle_knn / oracle_knn_bf
 *       the reference's golden vectors (tests/include/igl/knn.cpp: cube corners
 *       with a tie, and nv <= k), each against the other (best-first octree
 *       search vs brute force over the octree's candidates) on random and
 *       clustered sets, and coincident points (only the first is reachable).
 *   oracle_py.point_areas (Python: numpy SVD + scipy Delaunay)
 *       closed forms: the Voronoi area of a lattice point ((sqrt 3 / 2) a^2 on
 *       a triangular lattice, scaled by the stretch of a stretched one), the
 *       plane normal of planar neighbourhoods, the normal filter, and the
 *       degenerate cases (<= 2 neighbours, collinear neighbours: 0).
 *   oracle_octree / oracle_cloud_*
 *       the reference golden values of tests/include/igl/fast_winding_number.cpp:46-52
 *       (one-point cloud, 1e-15), the beta<=0 brute-force branch
 *       (fast_winding_number.cpp:300-310) at large beta, hand-checked tiny
 *       octrees (octree.cpp rules) and the closed-form Taylor terms.
 */
#ifndef GWN_ORACLE_H
#define GWN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scheduler: igl::parallel_for (parallel_for.h:140-204) ---- */
typedef void (*oracle_range_fn)(int64_t begin, int64_t end, void* ctx);
/* Runs fn over [0,n) in contiguous slices of round((n+1)/T) on T std::threads
 * (T = nthreads, or IGL_NUM_THREADS / hardware concurrency when nthreads<=0);
 * serial when n < min_parallel.  Returns the thread count used. */
int oracle_parallel_for(int64_t n, oracle_range_fn fn, void* ctx, int64_t min_parallel, int nthreads);
int oracle_default_num_threads(void);

/* ---- exact FP64 (solid_angle.cpp:10-55, winding_number.cpp:62-89) ---- */
double oracle_solid_angle(const double* a, const double* b, const double* c, const double* p);
/* brute-force sum over faces in F order; V nv x 3, F nf x 3 row-major; W[no] */
void oracle_winding_number_bf(const double* V, int64_t nv, const int32_t* F, int64_t nf,
                              const double* O, int64_t no, double* W, int nthreads);
/* FP64 restatement of UTsignedSolidAngleTri (FastWindingNumberForSoups.h:6070-6110) */
double oracle_solid_angle_tri_f64(const double* a, const double* b, const double* c, const double* q);

/* ---- triangle soup: UT_SolidAngle<float,float> over UT::BVH<4> ---- */
typedef struct oracle_soup oracle_soup;
/* UT_SolidAngle::init (FastWindingNumberForSoups.h:6543-7133) */
oracle_soup* oracle_soup_build(const float* V, int32_t nv, const int32_t* F, int32_t nf, int order);
/* The same over a given tree (nnodes x 4 uint32 in the reference's encoding). */
oracle_soup* oracle_soup_build_nodes(const float* V, int32_t nv, const int32_t* F, int32_t nf, int order,
                                     const uint32_t* nodes, int32_t nnodes);
void oracle_soup_free(oracle_soup* s);
int32_t oracle_soup_num_nodes(const oracle_soup* s);
/* node child arrays, 4 uint32 per node (FastWindingNumberForSoups.h:4217-4232) */
void oracle_soup_nodes(const oracle_soup* s, uint32_t* out);
/* BoxData, 92 floats per node in the reference struct order (6477-6521) */
void oracle_soup_boxdata(const oracle_soup* s, float* out);
/* computeSolidAngle (7148-7284) for one query; returns Omega (not /4pi) */
float oracle_soup_solid_angle(const oracle_soup* s, const float* q, float accuracy_scale);
/* fast_winding_number(bvh, accuracy_scale, Q, W) (fast_winding_number.cpp:418-437):
 * Q nq x 3 double narrowed to float, W = Omega / (4 pi) in double. */
void oracle_soup_query(const oracle_soup* s, const double* Q, int64_t nq, float accuracy_scale,
                       double* W, int nthreads);
/* Counted FLOPs / visits for the roofline: per query returns visits (node pre()
 * calls), accepted far-field child slots and exact triangles. */
void oracle_soup_count(const oracle_soup* s, const float* q, float accuracy_scale,
                       int64_t* visits, int64_t* far_accepted, int64_t* exact_tris);
/* The same, plus the exact triangles within near_d of the query's slot sphere
 * (|q - P| < sqrt(maxP2) + near_d): the mixed form's FP64 near terms. */
void oracle_soup_count_near(const oracle_soup* s, const float* q, float accuracy_scale, float near_d,
                            int64_t* visits, int64_t* far_accepted, int64_t* exact_tris, int64_t* near_tris);

/* BVH::init<BOX_AREA> (FastWindingNumberForSoups.h:4579-5944) over nboxes
 * boxes of 6 floats (x0 x1 y0 y1 z0 z1; 2D boxes carry z0 = z1 = 0 with
 * naxes = 2).  Returns the node count; *nodes (4 uint32 per node, malloc'ed,
 * free()d by the caller) is NULL for an empty tree. */
uint32_t oracle_bvh4_build(const float* boxes6, uint32_t nboxes, int naxes, uint32_t** nodes);

/* ---- 2D curve: UT_SubtendedAngle<float,float> over UT::BVH<4> ---- */
typedef struct oracle_curve oracle_curve;
/* UT_SubtendedAngle::init (FastWindingNumberForSoups.h:7342-7604): P np x 2,
 * E ne x 2 segment point indices */
oracle_curve* oracle_curve_build(const float* P, int32_t np, const int32_t* E, int32_t ne, int order);
void oracle_curve_free(oracle_curve* c);
int32_t oracle_curve_num_nodes(const oracle_curve* c);
void oracle_curve_nodes(const oracle_curve* c, uint32_t* out);
/* BoxData, 48 floats per node in the reference struct order (7287-7320) */
void oracle_curve_boxdata(const oracle_curve* c, float* out);
/* computeAngle (7700-7828) for each query (Q nq x 2 float): the signed
 * subtended angle (the winding number is angle / 2 pi) */
void oracle_curve_angle(const oracle_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle,
                        int nthreads);
/* Counted visits for the roofline: node pre() calls, accepted far-field child
 * slots and exact segments of one query. */
void oracle_curve_count(const oracle_curve* c, const float* q, float accuracy_scale, int64_t* visits,
                        int64_t* far_accepted, int64_t* exact_segments);

/* ---- signed distance (signed_distance.cpp:28-140) ---- */
/* closest point of triangle (a,b,c) to p, point_simplex_squared_distance.cpp:43-114;
 * writes it to out (may be NULL) and returns the squared distance */
double oracle_point_tri_sqrdist(const double* p, const double* a, const double* b, const double* c, double* out);
/* igl::signed_distance(P,V,F,sign_type,lower,upper,S,I,C,N) for sign types
 * 0 (pseudonormal), 1/2 (winding number), 3 (unsigned), 4 (fast winding
 * number), by brute force over all faces; returns 0, or -1 for an unknown
 * sign type */
int oracle_signed_distance(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* P, int64_t np,
                           int sign_type, double lower, double upper, double* S, int32_t* I, double* C,
                           int nthreads);
/* The same with N np x 3 (may be NULL): the pseudonormal sign type (0) writes
 * the normal it used (0 for points out of bounds). */
int oracle_signed_distance_n(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* P, int64_t np,
                             int sign_type, double lower, double upper, double* S, int32_t* I, double* C, double* N,
                             int nthreads);
/* per_face_normals, per_vertex_normals (angle weights) and per_edge_normals
 * (uniform weights) of (V,F): FN nf x 3, VN nv x 3, EN nf x 3 x 3 (the edge
 * opposite corner c of face f).  0 on success. */
int oracle_pseudonormals(const double* V, int64_t nv, const int32_t* F, int64_t nf, double* FN, double* VN,
                         double* EN);
/* pseudonormal_test: the normal for closest point c on face f. */
void oracle_pseudonormal_at(const double* V, const int32_t* F, const double* FN, const double* VN, const double* EN,
                            int64_t f, const double* c, double* n);

/* ---- 2D curves (winding_number.cpp:62-89 with F.cols() == 2) ----
 * signed_angle.cpp:8-41; V nv x 2, E ne x 2, O no x 2 */
double oracle_signed_angle_2d(const double* a, const double* b, const double* p);
void oracle_winding_number_2d(const double* V, const int32_t* E, int64_t ne, const double* O, int64_t no,
                              double* W);

/* ---- point cloud: igl::octree + fast_winding_number(P,N,A,...) ---- */
/* octree (octree.cpp:7-169).  Cells are numbered exactly as the reference
 * numbers them.  Outputs are malloc'ed and owned by the caller (free()).
 *   CH: ncells x 8 int32 (-1 for leaves)
 *   CN: ncells x 3, W: ncells (cell centres / widths, in REAL)
 *   cell_start/cell_count + point_list: the points of each cell in ascending
 *   index order (point_indices in the reference), as CSR. */
typedef struct {
  int32_t ncells;
  int32_t* CH;
  void* CN;
  void* W;
  int64_t* cell_start;
  int32_t* cell_count;
  int32_t* point_list;
  int64_t npoint_list;
} oracle_octree_t;

int oracle_octree_f64(const double* P, int32_t n, oracle_octree_t* out);
int oracle_octree_f32(const float* P, int32_t n, oracle_octree_t* out);
void oracle_octree_release(oracle_octree_t* t);

/* precompute (fast_winding_number.cpp:8-124): CM ncells x 3, R ncells,
 * EC ncells x num_terms (3 / 12 / 39 for order 0 / 1 / 2) */
int oracle_cloud_precompute_f64(const double* P, const double* N, const double* A,
                                const oracle_octree_t* t, int order,
                                double* CM, double* R, double* EC);
int oracle_cloud_precompute_f32(const float* P, const float* N, const float* A,
                                const oracle_octree_t* t, int order,
                                float* CM, float* R, float* EC);
/* evaluation (fast_winding_number.cpp:126-311); beta <= 0 is the brute-force
 * branch (300-310). */
void oracle_cloud_eval_f64(const double* P, const double* N, const double* A,
                           const oracle_octree_t* t, int order,
                           const double* CM, const double* R, const double* EC,
                           const double* Q, int64_t nq, double beta, double* WN, int nthreads);
void oracle_cloud_eval_f32(const float* P, const float* N, const float* A,
                           const oracle_octree_t* t, int order,
                           const float* CM, const float* R, const float* EC,
                           const float* Q, int64_t nq, double beta, float* WN, int nthreads);
/* Counted work of one query (bench.py's roofline): child-cell distance
 * tests, direct dipole terms, Taylor expansions. */
void oracle_cloud_count_f64(const oracle_octree_t* t, const double* CM, const double* R, const double* q,
                            double beta, int64_t* tests, int64_t* dipoles, int64_t* exps);
void oracle_cloud_count_f32(const oracle_octree_t* t, const float* CM, const float* R, const float* q,
                            double beta, int64_t* tests, int64_t* dipoles, int64_t* exps);

/* ---- antipodal method (WindingNumberAntipodalScene.h:48-256) ----
 * x0 NULL: the reference's default direction.  *nseg_out (may be NULL)
 * receives the number of weighted boundary segments. */
int oracle_winding_number_antipodal(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* x0,
                                    const double* O, int64_t no, double* W, int64_t* nseg_out, int nthreads);

/* ---- k nearest neighbours (knn.cpp:24-127) ----
 * oracle_knn: the reference's best-first search over an FP64 octree of V
 * (oracle_octree_f64); I np x min(nv, k).  Returns -1 when fewer than k points
 * are reachable.  oracle_knn_bf: brute force over cand[0 .. ncand) (NULL: all
 * of V), ordered by (squared distance, index). */
int oracle_knn(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const oracle_octree_t* t,
               int32_t* I, int nthreads);
int oracle_knn_bf(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand,
                  int64_t ncand, int32_t* I, int nthreads);

/* ---- exact winding number through the reference's hierarchy
 * (winding_number.cpp:39-53, WindingNumberAABB.h, WindingNumberTree.h) ----
 * oracle_wntree_build calls rand() in the reference's order.
 * oracle_wntree_nodes: 5 int64 per node in construction order: #faces,
 * #cap triangles, fan apex (-1 without exterior edges), left, right (-1). */
typedef struct oracle_wntree oracle_wntree;
oracle_wntree* oracle_wntree_build(const double* V, int64_t nv, const int32_t* F, int64_t nf);
void oracle_wntree_free(oracle_wntree* t);
int64_t oracle_wntree_num_nodes(const oracle_wntree* t);
void oracle_wntree_nodes(const oracle_wntree* t, int64_t* out);
void oracle_wntree_query(const oracle_wntree* t, const double* O, int64_t no, double* W, int nthreads);
/* the number of solid angles the query of p evaluates */
int64_t oracle_wntree_count(const oracle_wntree* t, const double* p);

#ifdef __cplusplus
}
#endif
#endif
