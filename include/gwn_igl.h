// gwn_igl.h -- libigl-signature facade over the gwn C ABI (header only).
//
// Replaces, with the same names, argument meaning and output resizing:
//   igl::winding_number(V,F,O,W)                      winding_number.h:41-48 (triangles, or 2D edges)
//   igl::winding_number(V,F,p)                        winding_number.h:49-60
//   igl::WindingNumberAABB<Scalar,Index>              WindingNumberAABB.h (the exact hierarchy)
//   igl::octree(P,point_indices,CH,CN,W)              octree.h:36-41
//   igl::fast_winding_number(P,N,A,point_indices,CH,expansion_order,CM,R,EC)
//                                                     fast_winding_number.h:44-53
//   igl::fast_winding_number(P,N,A,point_indices,CH,CM,R,EC,Q,beta,WN)
//                                                     fast_winding_number.h:88-99
//   igl::fast_winding_number(V,F,Q,W)                 fast_winding_number.h:160-165
//   igl::fast_winding_number(V,F,order,bvh)           fast_winding_number.h:176-180
//   igl::fast_winding_number(bvh,accuracy_scale,Q,W)  fast_winding_number.h:191-196
//   igl::fast_winding_number(bvh,accuracy_scale,p)    fast_winding_number.h:203-206
//   igl::fast_winding_number(P,N,A,Q,order,beta,WN)   fast_winding_number.h:111-118
//   igl::fast_winding_number(P,N,A,Q,WN)              fast_winding_number.h:126-131
//   struct igl::FastWindingNumberBVH                  fast_winding_number.h:140-152
//   igl::read_triangle_mesh(str,V,F)                   read_triangle_mesh.h (.off / .obj)
//   igl::FastWindingNumber::HDK_Sample::UT_SubtendedAngle<float,float>
//                                                     FastWindingNumberForSoups.h:6315-6370 (2D curves)
//   igl::winding_number_antipodal(V,F,intersector,O,W) winding_number_antipodal.h:43-53
//   igl::WindingNumberAntipodalScene<Scalar>          WindingNumberAntipodalScene.h:48-256
//   igl::knn(P,k,point_indices,CH,CN,W,I)             knn.h:33-45
//   igl::knn(P,V,k,point_indices,CH,CN,W,I)           knn.h:51-67
//   igl::copyleft::cgal::point_areas(P,I,N,A[,T])     copyleft/cgal/point_areas.h:44-62
//   igl::copyleft::cgal::fast_winding_number(P,N,Q[,expansion_order,beta],WN)
//                                                     copyleft/cgal/fast_winding_number.h:35-53
//   igl::SignedDistanceType                           signed_distance.h:20-34
//   igl::signed_distance(P,V,F,type,lower,upper,S,I,C,N)  signed_distance.h:61-71
//   igl::signed_distance(P,V,F,type,S,I,C,N)          signed_distance.h:81-89
//   igl::signed_distance_fast_winding_number(P,V,F,tree,fwn_bvh,S)
//                                                     signed_distance.h:276-289 (batch, .cpp:665-687)
//   igl::signed_distance_fast_winding_number(q,V,F,tree,fwn_bvh)
//                                                     signed_distance.h:300-311 (one point, .cpp:688-711)
//   igl::signed_distance_winding_number(tree,V,F,hier,q[,s,sqrd,i,c])
//                                                     signed_distance.h:211-242 (3D, .cpp:579-629)
// (signed distance for all five sign types: unsigned, winding number, fast
// winding number, pseudonormal, and the default, which is the winding-number
// type as in the reference; N is filled for the pseudonormal type)
// A program switches by including this header instead of the libigl ones and
// linking libgwn.so.  Matrices of any storage order and scalar type are
// accepted (converted to the row-major float/double buffers of the C ABI);
// outputs are resized by the callee as in the reference.  Where the reference
// asserts, these wrappers throw std::runtime_error carrying gwn_last_error().
//
// The conversions live in igl::gwn_detail as templates over any matrix type
// with rows(), cols(), operator()(i,j) and resize(r,c); the Eigen overloads
// are declared when <Eigen/Core> is available.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "gwn.h"

namespace igl {

enum SignedDistanceType {
  SIGNED_DISTANCE_TYPE_PSEUDONORMAL = GWN_SIGN_PSEUDONORMAL,
  SIGNED_DISTANCE_TYPE_WINDING_NUMBER = GWN_SIGN_WINDING_NUMBER,
  SIGNED_DISTANCE_TYPE_DEFAULT = GWN_SIGN_DEFAULT,
  SIGNED_DISTANCE_TYPE_UNSIGNED = GWN_SIGN_UNSIGNED,
  SIGNED_DISTANCE_TYPE_FAST_WINDING_NUMBER = GWN_SIGN_FAST_WINDING_NUMBER,
  NUM_SIGNED_DISTANCE_TYPE = 5
};

// igl::FastWindingNumberBVH: owns the GPU tree.  Like the reference's (which
// holds raw pointers into itself) it is neither copyable nor movable.
struct FastWindingNumberBVH {
  gwn_soup* handle = nullptr;
  int tree_kind = GWN_TREE_SAH4;   // reference topology by default
  int precision = GWN_PREC_FP32;   // reference arithmetic by default
  // fingerprint of the vertices the tree was built from (as doubles), so that
  // signed_distance_fast_winding_number can tell whether its closest points
  // may come from this tree's vertices
  uint64_t vertex_hash = 0;
  int64_t num_vertices = -1;
  FastWindingNumberBVH() = default;
  FastWindingNumberBVH(const FastWindingNumberBVH&) = delete;
  FastWindingNumberBVH& operator=(const FastWindingNumberBVH&) = delete;
  ~FastWindingNumberBVH() {
    if (handle) gwn_soup_free(handle);
  }
};

namespace gwn_detail {

inline void check(int rc) {
  if (rc != GWN_OK) throw std::runtime_error(std::string("gwn: ") + gwn_last_error());
}

// FNV-1a over the bytes of a double array
inline uint64_t hash_doubles(const std::vector<double>& v) {
  uint64_t h = 1469598103934665603ull;
  const unsigned char* b = reinterpret_cast<const unsigned char*>(v.data());
  for (size_t i = 0; i < v.size() * sizeof(double); ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

template <typename T, typename M>
std::vector<T> rows3(const M& X, const char* what) {
  if (X.rows() > 0 && X.cols() != 3) throw std::runtime_error(std::string(what) + " should have 3 columns");
  std::vector<T> out(3 * (size_t)X.rows());
  for (int64_t i = 0; i < (int64_t)X.rows(); ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = static_cast<T>(X(i, j));
  return out;
}

template <typename M>
std::vector<int32_t> faces(const M& F) {
  if (F.rows() > 0 && F.cols() != 3) throw std::runtime_error("F should contain triangles");
  std::vector<int32_t> out(3 * (size_t)F.rows());
  for (int64_t i = 0; i < (int64_t)F.rows(); ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = static_cast<int32_t>(F(i, j));
  return out;
}

}  // namespace gwn_detail

// igl::WindingNumberAABB<Scalar, Index> (WindingNumberAABB.h): the hierarchy
// igl::winding_number(V,F,O,W) builds, kept on the GPU.  Construct it over
// (V, F), call grow() (which builds it, drawing rand() as the reference's
// constructors and grow() do), then winding_number(p).  Without grow() the
// reference keeps only the root; here the first query builds the tree (call
// grow() before querying from several threads).  Not copyable; freed with
// the object.
template <typename Scalar, typename Index>
class WindingNumberAABB;
template <typename T>
struct is_winding_number_aabb : std::false_type {};
template <typename S, typename I>
struct is_winding_number_aabb<WindingNumberAABB<S, I>> : std::true_type {};

template <typename Scalar, typename Index>
class WindingNumberAABB {
 public:
  WindingNumberAABB() = default;
  template <typename MV, typename MF>
  WindingNumberAABB(const MV& V, const MF& F) {
    set_mesh(V, F);
  }
  WindingNumberAABB(const WindingNumberAABB&) = delete;
  WindingNumberAABB& operator=(const WindingNumberAABB&) = delete;
  ~WindingNumberAABB() {
    if (handle) gwn_exact_tree_free(handle);
  }
  template <typename MV, typename MF>
  void set_mesh(const MV& V, const MF& F) {
    if (handle) gwn_exact_tree_free(handle);
    handle = nullptr;
    V_ = gwn_detail::rows3<double>(V, "V");
    F_ = gwn_detail::faces(F);
  }
  void grow() {
    if (handle) gwn_exact_tree_free(handle);
    handle = nullptr;
    gwn_detail::check(gwn_exact_tree_build(V_.data(), (int64_t)V_.size() / 3, F_.data(), (int64_t)F_.size() / 3,
                                           &handle));
  }
  // WindingNumberTree::winding_number(p) (WindingNumberTree.h:262-336)
  template <typename Mp>
  Scalar winding_number(const Mp& p) const {
    const double o[3] = {static_cast<double>(p(0, 0)), static_cast<double>(p(0, 1)), static_cast<double>(p(0, 2))};
    double w = 0.0;
    gwn_detail::check(gwn_exact_tree_query(tree(), o, 1, &w));
    return static_cast<Scalar>(w);
  }
  // the built tree (built here if grow() was not called)
  const gwn_exact_tree* tree() const {
    if (!handle) const_cast<WindingNumberAABB*>(this)->grow();
    return handle;
  }
  gwn_exact_tree* handle = nullptr;

 private:
  std::vector<double> V_;
  std::vector<int32_t> F_;
};

namespace gwn_detail {

template <typename MV, typename MF, typename MO, typename MW>
void winding_number(const MV& V, const MF& F, const MO& O, MW& W) {
  if (F.cols() == 2) {  // 2D curve of edges (winding_number.cpp:24-33)
    if (V.cols() != 2 || (O.rows() > 0 && O.cols() != 2)) throw std::runtime_error("2D winding number needs 2D V and O");
    std::vector<double> v(2 * (size_t)V.rows()), o(2 * (size_t)O.rows());
    std::vector<int32_t> e(2 * (size_t)F.rows());
    for (int64_t i = 0; i < (int64_t)V.rows(); ++i)
      for (int j = 0; j < 2; ++j) v[2 * i + j] = static_cast<double>(V(i, j));
    for (int64_t i = 0; i < (int64_t)F.rows(); ++i)
      for (int j = 0; j < 2; ++j) e[2 * i + j] = static_cast<int32_t>(F(i, j));
    for (int64_t i = 0; i < (int64_t)O.rows(); ++i)
      for (int j = 0; j < 2; ++j) o[2 * i + j] = static_cast<double>(O(i, j));
    std::vector<double> w((size_t)O.rows());
    if (O.rows() > 0)
      check(gwn_exact_winding_number_2d(v.data(), V.rows(), e.data(), F.rows(), o.data(), O.rows(), w.data()));
    W.resize(O.rows(), 1);
    for (int64_t i = 0; i < (int64_t)O.rows(); ++i) W(i, 0) = static_cast<typename std::decay<decltype(W(0, 0))>::type>(w[i]);
    return;
  }
  const auto v = rows3<double>(V, "V");
  const auto f = faces(F);
  const auto o = rows3<double>(O, "O");
  std::vector<double> w((size_t)O.rows());
  check(gwn_exact_winding_number(v.data(), V.rows(), f.data(), F.rows(), o.data(), O.rows(), w.data()));
  W.resize(O.rows(), 1);
  for (int64_t i = 0; i < (int64_t)O.rows(); ++i) W(i, 0) = static_cast<typename std::decay<decltype(W(0, 0))>::type>(w[i]);
}

template <typename MV, typename MF, typename Mp>
double winding_number_point(const MV& V, const MF& F, const Mp& p) {
  const auto v = rows3<double>(V, "V");
  const auto f = faces(F);
  const double o[3] = {static_cast<double>(p(0, 0)), static_cast<double>(p(0, 1)), static_cast<double>(p(0, 2))};
  double w = 0;  // winding_number.cpp:62-89: the sum over every face
  check(gwn_exact_winding_number_bruteforce(v.data(), V.rows(), f.data(), F.rows(), o, 1, &w));
  return w;
}

template <typename MV, typename MF>
void build(const MV& V, const MF& F, int order, FastWindingNumberBVH& bvh) {
  const auto v = rows3<double>(V, "V");
  const auto f = faces(F);
  if (bvh.handle) {
    gwn_soup_free(bvh.handle);
    bvh.handle = nullptr;
  }
  check(gwn_soup_build(v.data(), V.rows(), f.data(), F.rows(), order, bvh.tree_kind, &bvh.handle));
  bvh.vertex_hash = hash_doubles(v);
  bvh.num_vertices = V.rows();
}

template <typename MQ, typename MW>
void query(const FastWindingNumberBVH& bvh, float accuracy_scale, const MQ& Q, MW& W) {
  if (!bvh.handle) throw std::runtime_error("FastWindingNumberBVH not built");
  const auto q = rows3<double>(Q, "Q");
  std::vector<double> w((size_t)Q.rows());
  if (Q.rows() > 0) check(gwn_soup_query(bvh.handle, q.data(), Q.rows(), accuracy_scale, bvh.precision, w.data()));
  W.resize(Q.rows(), 1);
  for (int64_t i = 0; i < (int64_t)Q.rows(); ++i) W(i, 0) = static_cast<typename std::decay<decltype(W(0, 0))>::type>(w[i]);
}

template <typename Mp>
double query_point(const FastWindingNumberBVH& bvh, float accuracy_scale, const Mp& p) {
  if (!bvh.handle) throw std::runtime_error("FastWindingNumberBVH not built");
  const double q[3] = {static_cast<double>(p(0, 0)), static_cast<double>(p(0, 1)), static_cast<double>(p(0, 2))};
  double w = 0;
  check(gwn_soup_query(bvh.handle, q, 1, accuracy_scale, bvh.precision, &w));
  return w;
}

template <typename MP, typename MN, typename MA, typename MQ, typename MW>
void cloud(const MP& P, const MN& N, const MA& A, const MQ& Q, int order, double beta, MW& WN) {
  const auto p = rows3<float>(P, "P");
  const auto n = rows3<float>(N, "N");
  std::vector<float> a((size_t)P.rows());
  for (int64_t i = 0; i < (int64_t)P.rows(); ++i) a[i] = static_cast<float>(A(i, 0));
  const auto q = rows3<float>(Q, "Q");
  gwn_cloud* c = nullptr;
  check(gwn_cloud_build(p.data(), n.data(), a.data(), P.rows(), order, &c));
  std::vector<float> w((size_t)Q.rows());
  const int rc = Q.rows() > 0 ? gwn_cloud_query(c, q.data(), Q.rows(), beta, w.data()) : GWN_OK;
  gwn_cloud_free(c);
  check(rc);
  WN.resize(Q.rows(), 1);
  for (int64_t i = 0; i < (int64_t)Q.rows(); ++i) WN(i, 0) = static_cast<typename std::decay<decltype(WN(0, 0))>::type>(w[i]);
}

// point_indices (vector of vectors) -> cell_start (ncells + 1) / point_list
template <typename Index>
void flatten_lists(const std::vector<std::vector<Index>>& point_indices, std::vector<int64_t>& start,
                   std::vector<int32_t>& list) {
  start.assign(point_indices.size() + 1, 0);
  for (size_t i = 0; i < point_indices.size(); ++i) start[i + 1] = start[i] + (int64_t)point_indices[i].size();
  list.clear();
  list.reserve((size_t)start.back());
  for (const auto& cell : point_indices)
    for (const auto& k : cell) list.push_back(static_cast<int32_t>(k));
}

template <typename MCH>
std::vector<int32_t> children(const MCH& CH) {
  if (CH.rows() > 0 && CH.cols() != 8) throw std::runtime_error("CH should have 8 columns");
  std::vector<int32_t> out(8 * (size_t)CH.rows());
  for (int64_t i = 0; i < (int64_t)CH.rows(); ++i)
    for (int j = 0; j < 8; ++j) out[8 * i + j] = static_cast<int32_t>(CH(i, j));
  return out;
}

template <typename MP, typename Index, typename MCH, typename MCN, typename MW>
void octree(const MP& P, std::vector<std::vector<Index>>& point_indices, MCH& CH, MCN& CN, MW& W) {
  const auto p = rows3<float>(P, "P");
  gwn_cloud* c = nullptr;
  check(gwn_cloud_build(p.data(), nullptr, nullptr, P.rows(), 0, &c));
  int64_t m = 0;
  std::vector<int32_t> ch, list;
  std::vector<float> cn, w;
  std::vector<int64_t> cs;
  std::vector<int32_t> cc;
  int rc = gwn_cloud_num_cells(c, &m);
  if (rc == GWN_OK) {
    ch.resize(8 * (size_t)m);
    cn.resize(3 * (size_t)m);
    w.resize((size_t)m);
    cs.resize((size_t)m);
    cc.resize((size_t)m);
    list.resize((size_t)std::max<int64_t>(1, P.rows()));
    rc = gwn_cloud_copy_octree(c, ch.data(), cn.data(), w.data(), cs.data(), cc.data(), list.data());
  }
  gwn_cloud_free(c);
  check(rc);
  // octree.cpp lists each cell's points in ascending index order
  point_indices.assign((size_t)m, {});
  for (int64_t i = 0; i < m; ++i) {
    auto& cell = point_indices[i];
    cell.reserve((size_t)cc[i]);
    for (int32_t k = 0; k < cc[i]; ++k) cell.push_back(static_cast<Index>(list[cs[i] + k]));
    std::sort(cell.begin(), cell.end());
  }
  CH.resize(m, 8);
  CN.resize(m, 3);
  W.resize(m, 1);
  for (int64_t i = 0; i < m; ++i) {
    for (int j = 0; j < 8; ++j) CH(i, j) = ch[8 * i + j];
    for (int j = 0; j < 3; ++j) CN(i, j) = cn[3 * i + j];
    W(i, 0) = w[i];
  }
}

template <typename MP, typename MN, typename MA, typename Index, typename MCH>
gwn_cloud* cloud_over(const MP& P, const MN& N, const MA& A, const std::vector<std::vector<Index>>& point_indices,
                      const MCH& CH, int order, const float* CM, const float* R, const float* EC) {
  const auto p = rows3<float>(P, "P");
  const auto n = rows3<float>(N, "N");
  std::vector<float> a((size_t)P.rows());
  for (int64_t i = 0; i < (int64_t)P.rows(); ++i) a[i] = static_cast<float>(A(i, 0));
  const auto ch = children(CH);
  if ((int64_t)point_indices.size() != (int64_t)CH.rows())
    throw std::runtime_error("point_indices and CH disagree on the number of cells");
  std::vector<int64_t> start;
  std::vector<int32_t> list;
  flatten_lists(point_indices, start, list);
  gwn_cloud* c = nullptr;
  check(gwn_cloud_build_from_octree(p.data(), n.data(), a.data(), P.rows(), ch.data(), CH.rows(), start.data(),
                                    list.data(), order, CM, R, EC, &c));
  return c;
}

template <typename MP, typename MN, typename MA, typename Index, typename MCH, typename MCM, typename MR,
          typename MEC>
void cloud_precompute(const MP& P, const MN& N, const MA& A, const std::vector<std::vector<Index>>& point_indices,
                      const MCH& CH, int order, MCM& CM, MR& R, MEC& EC) {
  if (order < 0 || order > 2) throw std::runtime_error("expansion_order must be 0, 1 or 2");
  gwn_cloud* c = cloud_over(P, N, A, point_indices, CH, order, nullptr, nullptr, nullptr);
  const int64_t m = CH.rows();
  const int nt = order == 0 ? 3 : (order == 1 ? 12 : 39);
  std::vector<float> cm(3 * (size_t)m), r((size_t)m), ec((size_t)nt * m);
  const int rc = m > 0 ? gwn_cloud_copy_expansion(c, cm.data(), r.data(), ec.data()) : GWN_OK;
  gwn_cloud_free(c);
  check(rc);
  CM.resize(m, 3);
  R.resize(m, 1);
  EC.resize(m, nt);
  for (int64_t i = 0; i < m; ++i) {
    for (int j = 0; j < 3; ++j) CM(i, j) = cm[3 * i + j];
    R(i, 0) = r[i];
    for (int j = 0; j < nt; ++j) EC(i, j) = ec[(size_t)nt * i + j];
  }
}

template <typename MP, typename MN, typename MA, typename Index, typename MCH, typename MCM, typename MR,
          typename MEC, typename MQ, typename MW>
void cloud_eval(const MP& P, const MN& N, const MA& A, const std::vector<std::vector<Index>>& point_indices,
                const MCH& CH, const MCM& CM, const MR& R, const MEC& EC, const MQ& Q, double beta, MW& WN) {
  const int64_t m = CH.rows();
  const int nt = (int)EC.cols();
  if (nt != 3 && nt != 12 && nt != 39) throw std::runtime_error("EC should have 3, 12 or 39 columns");
  if ((int64_t)CM.rows() != m || (int64_t)R.rows() * (int64_t)R.cols() != m || (int64_t)EC.rows() != m)
    throw std::runtime_error("CM, R, EC and CH disagree on the number of cells");
  const auto cm = rows3<float>(CM, "CM");
  std::vector<float> r((size_t)m), ec((size_t)nt * m);
  for (int64_t i = 0; i < m; ++i) {
    r[i] = static_cast<float>(R(i, 0));
    for (int j = 0; j < nt; ++j) ec[(size_t)nt * i + j] = static_cast<float>(EC(i, j));
  }
  const int order = nt == 3 ? 0 : (nt == 12 ? 1 : 2);
  gwn_cloud* c = cloud_over(P, N, A, point_indices, CH, order, cm.data(), r.data(), ec.data());
  const auto q = rows3<float>(Q, "Q");
  std::vector<float> w((size_t)Q.rows());
  const int rc = Q.rows() > 0 ? gwn_cloud_query(c, q.data(), Q.rows(), beta, w.data()) : GWN_OK;
  gwn_cloud_free(c);
  check(rc);
  WN.resize(Q.rows(), 1);
  for (int64_t i = 0; i < (int64_t)Q.rows(); ++i)
    WN(i, 0) = static_cast<typename std::decay<decltype(WN(0, 0))>::type>(w[i]);
}

// signed distance over a built soup; S, I, C resized as the reference does,
// and N (np x 3) for the pseudonormal type only (signed_distance.cpp:68-76)
template <typename MP, typename MS, typename MI, typename MC, typename MN>
void signed_distance_bvh(gwn_soup* h, const MP& P, int type, double lower, double upper, MS& S, MI& I, MC& C,
                         MN* N) {
  const auto p = rows3<double>(P, "P");
  const int64_t n = P.rows();
  const bool want_n = N != nullptr && type == GWN_SIGN_PSEUDONORMAL;
  std::vector<double> s((size_t)n), c(3 * (size_t)n), nn(want_n ? 3 * (size_t)n : 0);
  std::vector<int32_t> i((size_t)n);
  if (n > 0)
    check(gwn_signed_distance_with_normals(h, p.data(), n, type, lower, upper, s.data(), i.data(), c.data(),
                                           want_n ? nn.data() : nullptr));
  S.resize(n, 1);
  I.resize(n, 1);
  C.resize(n, 3);
  for (int64_t k = 0; k < n; ++k) {
    S(k, 0) = static_cast<typename std::decay<decltype(S(0, 0))>::type>(s[k]);
    I(k, 0) = static_cast<typename std::decay<decltype(I(0, 0))>::type>(i[k]);
    for (int j = 0; j < 3; ++j) C(k, j) = static_cast<typename std::decay<decltype(C(0, 0))>::type>(c[3 * k + j]);
  }
  if (want_n) {
    N->resize(n, 3);
    for (int64_t k = 0; k < n; ++k)
      for (int j = 0; j < 3; ++j)
        (*N)(k, j) = static_cast<typename std::decay<decltype((*N)(0, 0))>::type>(nn[3 * k + j]);
  }
}

template <typename MP, typename MV, typename MF, typename MS, typename MI, typename MC, typename MN>
void signed_distance(const MP& P, const MV& V, const MF& F, int type, double lower, double upper, MS& S, MI& I,
                     MC& C, MN* N) {
  FastWindingNumberBVH bvh;  // signed_distance.cpp:65-67 builds an order-2 tree for the FWN sign
  build(V, F, 2, bvh);
  signed_distance_bvh(bvh.handle, P, type, lower, upper, S, I, C, N);
}

template <typename MP, typename MV, typename MF, typename MS, typename MI, typename MC>
void signed_distance(const MP& P, const MV& V, const MF& F, int type, double lower, double upper, MS& S, MI& I,
                     MC& C) {
  signed_distance(P, V, F, type, lower, upper, S, I, C, static_cast<MC*>(nullptr));
}

// The points igl::knn can return over an octree (knn.cpp:101-123): the first
// point of every non-empty leaf reachable from cell 0.
template <typename Index, typename MCH>
std::vector<int32_t> octree_candidates(const std::vector<std::vector<Index>>& point_indices, const MCH& CH) {
  std::vector<int32_t> out;
  if (CH.rows() == 0) return out;
  if (CH.cols() != 8) throw std::runtime_error("CH should have 8 columns");
  if ((int64_t)point_indices.size() != (int64_t)CH.rows())
    throw std::runtime_error("point_indices and CH disagree on the number of cells");
  std::vector<int64_t> stack(1, 0);
  std::vector<char> seen((size_t)CH.rows(), 0);
  while (!stack.empty()) {
    const int64_t c = stack.back();
    stack.pop_back();
    if (c < 0 || c >= (int64_t)CH.rows() || seen[c]) throw std::runtime_error("CH is not a tree");
    seen[c] = 1;
    if (CH(c, 0) == -1) {
      if (!point_indices[c].empty()) out.push_back(static_cast<int32_t>(point_indices[c][0]));
    } else {
      for (int j = 0; j < 8; ++j) stack.push_back(static_cast<int64_t>(CH(c, j)));
    }
  }
  std::sort(out.begin(), out.end());
  return out;
}

// knn.cpp:24-127: I (np x min(nv, k)) over the candidates of the caller's octree
template <typename MP, typename MV, typename Index, typename MCH, typename MI>
void knn(const MP& P, const MV& V, size_t k, const std::vector<std::vector<Index>>& point_indices, const MCH& CH,
         MI& I) {
  const auto p = rows3<double>(P, "P");
  const auto v = rows3<double>(V, "V");
  const int64_t np = P.rows(), nv = V.rows();
  const int cols = gwn_knn_columns(nv, (int64_t)k);
  std::vector<int32_t> out((size_t)np * cols);
  if (np > 0 && nv > 0) {
    const auto cand = octree_candidates(point_indices, CH);
    check(gwn_knn(p.data(), np, v.data(), nv, (int64_t)k, cand.data(), (int64_t)cand.size(), out.data()));
  }
  I.resize(np, cols);
  for (int64_t i = 0; i < np; ++i)
    for (int j = 0; j < cols; ++j)
      I(i, j) = static_cast<typename std::decay<decltype(I(0, 0))>::type>(out[(size_t)i * cols + j]);
}

// point_areas.cpp:37-170: A n x 1, T n x 3 (zero where no area is found)
template <typename MP, typename MI, typename MN, typename MA, typename MT>
void point_areas(const MP& P, const MI& I, const MN& N, MA& A, MT& T) {
  const auto p = rows3<double>(P, "P");
  const int64_t n = P.rows(), k = I.cols();
  if (I.rows() != n) throw std::runtime_error("P and I must have the same number of rows");
  std::vector<double> nn;
  if (N.rows() > 0) {
    if (N.rows() != n) throw std::runtime_error("P and N must have the same number of rows");
    nn = rows3<double>(N, "N");
  }
  std::vector<int32_t> idx((size_t)n * k);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < k; ++j) idx[(size_t)i * k + j] = static_cast<int32_t>(I(i, j));
  std::vector<double> a((size_t)n, 0.0), t(3 * (size_t)n, 0.0);
  if (n > 0 && k > 0)
    check(gwn_point_areas(p.data(), n, idx.data(), k, nn.empty() ? nullptr : nn.data(), a.data(), t.data()));
  A.resize(n, 1);
  T.resize(n, 3);
  for (int64_t i = 0; i < n; ++i) {
    A(i, 0) = static_cast<typename std::decay<decltype(A(0, 0))>::type>(a[i]);
    for (int j = 0; j < 3; ++j) T(i, j) = static_cast<typename std::decay<decltype(T(0, 0))>::type>(t[3 * i + j]);
  }
}

// copyleft/cgal/fast_winding_number.cpp:16-43: a cloud without areas
template <typename MP, typename MN, typename MQ, typename MW>
void cloud_estimated_areas(const MP& P, const MN& N, const MQ& Q, int order, double beta, MW& WN) {
  const auto p = rows3<double>(P, "P");
  const auto n = rows3<double>(N, "N");
  if (N.rows() != P.rows()) throw std::runtime_error("P and N must have the same number of rows");
  const auto q = rows3<float>(Q, "Q");
  gwn_cloud* c = nullptr;
  check(gwn_cloud_build_estimated_areas(p.data(), n.data(), P.rows(), order, nullptr, &c));
  std::vector<float> w((size_t)Q.rows());
  const int rc = Q.rows() > 0 ? gwn_cloud_query(c, q.data(), Q.rows(), beta, w.data()) : GWN_OK;
  gwn_cloud_free(c);
  check(rc);
  WN.resize(Q.rows(), 1);
  for (int64_t i = 0; i < (int64_t)Q.rows(); ++i)
    WN(i, 0) = static_cast<typename std::decay<decltype(WN(0, 0))>::type>(w[i]);
}

// signed_distance.cpp:611-629 for one point q: closest point on the soup `h`
// (squared distance recomputed from it, summed as the device sums it) and
// s = 1 - 2 w with the exact winding number, from the hierarchy `hier` when
// the caller passed one (hier.winding_number(q), as the reference), else
// summed over every face
template <typename MV, typename MF, typename Mq>
void signed_distance_wn_point(gwn_soup* h, const gwn_exact_tree* hier, const MV& V, const MF& F, const Mq& q,
                              double& s, double& sqrd, int& i, double c[3]) {
  const double p[3] = {static_cast<double>(q(0, 0)), static_cast<double>(q(0, 1)), static_cast<double>(q(0, 2))};
  double S = 0.0;
  int32_t I = -1;
  check(gwn_signed_distance(h, p, 1, GWN_SIGN_UNSIGNED, std::numeric_limits<double>::min(),
                            std::numeric_limits<double>::max(), &S, &I, c));
  const double e0 = p[0] - c[0], e1 = p[1] - c[1], e2 = p[2] - c[2];
  sqrd = (e0 * e0 + e1 * e1) + e2 * e2;
  i = I;
  double w = 0.0;
  if (hier) {
    check(gwn_exact_tree_query(hier, p, 1, &w));
  } else {  // the sum over every face (no hierarchy built for one point)
    const auto v = rows3<double>(V, "V");
    const auto f = faces(F);
    check(gwn_exact_winding_number_bruteforce(v.data(), V.rows(), f.data(), F.rows(), p, 1, &w));
  }
  s = 1. - 2. * w;
}

// WindingNumberAntipodalScene.h:69-78: the scene (x0 NULL: the default direction)
template <typename MV, typename MF>
gwn_antipodal* antipodal_build(const MV& V, const MF& F, const double* x0) {
  const auto v = rows3<double>(V, "V");
  const auto f = faces(F);
  gwn_antipodal* a = nullptr;
  check(gwn_antipodal_build(v.data(), V.rows(), f.data(), F.rows(), x0, &a));
  return a;
}

template <typename MO, typename MW>
void antipodal_query(const gwn_antipodal* a, const MO& O, MW& W) {
  const auto o = rows3<double>(O, "O");
  std::vector<double> w((size_t)O.rows());
  if (O.rows() > 0) check(gwn_antipodal_query(a, o.data(), O.rows(), w.data()));
  W.resize(O.rows(), 1);
  for (int64_t i = 0; i < (int64_t)O.rows(); ++i)
    W(i, 0) = static_cast<typename std::decay<decltype(W(0, 0))>::type>(w[i]);
}

// .off / .obj file; returns false (and leaves gwn_last_error) on failure, as
// the reference returns false
template <typename MV, typename MF>
bool read_triangle_mesh(const std::string& path, MV& V, MF& F) {
  double* v = nullptr;
  int32_t* f = nullptr;
  int64_t nv = 0, nf = 0;
  if (gwn_read_triangle_mesh(path.c_str(), &v, &nv, &f, &nf) != GWN_OK) return false;
  V.resize(nv, 3);
  F.resize(nf, 3);
  for (int64_t i = 0; i < nv; ++i)
    for (int j = 0; j < 3; ++j) V(i, j) = static_cast<typename std::decay<decltype(V(0, 0))>::type>(v[3 * i + j]);
  for (int64_t i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j) F(i, j) = static_cast<typename std::decay<decltype(F(0, 0))>::type>(f[3 * i + j]);
  gwn_free(v);
  gwn_free(f);
  return true;
}

}  // namespace gwn_detail

// FastWindingNumberForSoups.h's 2D fast winding number, UT_SubtendedAngle
// (6315-6370), in the reference's namespace.  Only <float,float> is
// instantiated, as in the reference.  Unlike the reference, the object copies
// the segments and positions to the GPU at init, so the caller's arrays need
// not outlive it.
namespace FastWindingNumber {
namespace HDK_Sample {

// the reference's UT_Vector2T is UT_FixedVector<T,2>: two components with []
template <typename T>
struct UT_Vector2T {
  T vec[2];
  UT_Vector2T() = default;
  UT_Vector2T(T x, T y) : vec{x, y} {}
  T& operator[](int i) { return vec[i]; }
  const T& operator[](int i) const { return vec[i]; }
};

template <typename T, typename S>
class UT_SubtendedAngle {
  static_assert(std::is_same<T, float>::value && std::is_same<S, float>::value,
                "UT_SubtendedAngle is provided for <float,float>");

 public:
  UT_SubtendedAngle() = default;
  UT_SubtendedAngle(const int nsegments, const int* const segment_points, const int npoints,
                    const UT_Vector2T<S>* const positions, const int order = 2) {
    init(nsegments, segment_points, npoints, positions, order);
  }
  UT_SubtendedAngle(const UT_SubtendedAngle&) = delete;
  UT_SubtendedAngle& operator=(const UT_SubtendedAngle&) = delete;
  ~UT_SubtendedAngle() { clear(); }

  // init (7342-7604); may be called again to re-initialise
  void init(const int nsegments, const int* const segment_points, const int npoints,
            const UT_Vector2T<S>* const positions, const int order = 2) {
    clear();
    std::vector<float> P(2 * (size_t)std::max(npoints, 0));
    for (int i = 0; i < npoints; ++i) {
      P[2 * (size_t)i] = positions[i][0];
      P[2 * (size_t)i + 1] = positions[i][1];
    }
    gwn_detail::check(gwn_curve_build(P.data(), npoints, segment_points, nsegments, order, GWN_TREE_SAH4, &handle_));
    nsegments_ = nsegments;
  }
  void clear() {
    if (handle_) gwn_curve_free(handle_);
    handle_ = nullptr;
    nsegments_ = 0;
  }
  bool isClear() const { return nsegments_ == 0; }

  // computeAngle (7700-7828): the signed angle the curve subtends at the point
  T computeAngle(const UT_Vector2T<T>& query_point, const T accuracy_scale = T(2.0)) const {
    T out = T(0);
    computeAngles(1, &query_point, &out, accuracy_scale);
    return out;
  }
  // batch form (no counterpart in the reference, which is called per point):
  // one GPU call for nq points
  void computeAngles(int64_t nq, const UT_Vector2T<T>* query_points, T* angles,
                     const T accuracy_scale = T(2.0)) const {
    if (nq <= 0) return;
    if (!handle_) {
      std::fill(angles, angles + nq, T(0));
      return;
    }
    static_assert(sizeof(UT_Vector2T<T>) == 2 * sizeof(T), "packed 2-vectors");
    gwn_detail::check(
        gwn_curve_query(handle_, reinterpret_cast<const float*>(query_points), nq, accuracy_scale, angles));
  }

 private:
  gwn_curve* handle_ = nullptr;
  int nsegments_ = 0;
};

}  // namespace HDK_Sample
}  // namespace FastWindingNumber
}  // namespace igl

#if defined(__has_include)
#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
namespace igl {

template <typename DerivedV, typename DerivedF, typename DerivedO, typename DerivedW>
void winding_number(const Eigen::MatrixBase<DerivedV>& V, const Eigen::MatrixBase<DerivedF>& F,
                    const Eigen::MatrixBase<DerivedO>& O, Eigen::PlainObjectBase<DerivedW>& W) {
  gwn_detail::winding_number(V.derived(), F.derived(), O.derived(), W.derived());
}

template <typename DerivedV, typename DerivedF, typename Derivedp>
typename DerivedV::Scalar winding_number(const Eigen::MatrixBase<DerivedV>& V, const Eigen::MatrixBase<DerivedF>& F,
                                         const Eigen::MatrixBase<Derivedp>& p) {
  return static_cast<typename DerivedV::Scalar>(gwn_detail::winding_number_point(V.derived(), F.derived(), p.derived()));
}

template <typename DerivedP, typename IndexType, typename DerivedCH, typename DerivedCN, typename DerivedW>
void octree(const Eigen::MatrixBase<DerivedP>& P, std::vector<std::vector<IndexType>>& point_indices,
            Eigen::PlainObjectBase<DerivedCH>& CH, Eigen::PlainObjectBase<DerivedCN>& CN,
            Eigen::PlainObjectBase<DerivedW>& W) {
  gwn_detail::octree(P.derived(), point_indices, CH.derived(), CN.derived(), W.derived());
}

template <typename DerivedP, typename DerivedA, typename DerivedN, typename Index, typename DerivedCH,
          typename DerivedCM, typename DerivedR, typename DerivedEC>
void fast_winding_number(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedN>& N,
                         const Eigen::MatrixBase<DerivedA>& A, const std::vector<std::vector<Index>>& point_indices,
                         const Eigen::MatrixBase<DerivedCH>& CH, const int expansion_order,
                         Eigen::PlainObjectBase<DerivedCM>& CM, Eigen::PlainObjectBase<DerivedR>& R,
                         Eigen::PlainObjectBase<DerivedEC>& EC) {
  gwn_detail::cloud_precompute(P.derived(), N.derived(), A.derived(), point_indices, CH.derived(), expansion_order,
                               CM.derived(), R.derived(), EC.derived());
}

template <typename DerivedP, typename DerivedA, typename DerivedN, typename Index, typename DerivedCH,
          typename DerivedCM, typename DerivedR, typename DerivedEC, typename DerivedQ, typename BetaType,
          typename DerivedWN>
void fast_winding_number(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedN>& N,
                         const Eigen::MatrixBase<DerivedA>& A, const std::vector<std::vector<Index>>& point_indices,
                         const Eigen::MatrixBase<DerivedCH>& CH, const Eigen::MatrixBase<DerivedCM>& CM,
                         const Eigen::MatrixBase<DerivedR>& R, const Eigen::MatrixBase<DerivedEC>& EC,
                         const Eigen::MatrixBase<DerivedQ>& Q, const BetaType beta,
                         Eigen::PlainObjectBase<DerivedWN>& WN) {
  gwn_detail::cloud_eval(P.derived(), N.derived(), A.derived(), point_indices, CH.derived(), CM.derived(),
                         R.derived(), EC.derived(), Q.derived(), (double)beta, WN.derived());
}

template <typename DerivedV, typename DerivedF>
bool read_triangle_mesh(const std::string str, Eigen::PlainObjectBase<DerivedV>& V,
                        Eigen::PlainObjectBase<DerivedF>& F) {
  return gwn_detail::read_triangle_mesh(str, V.derived(), F.derived());
}

template <typename DerivedP, typename DerivedV, typename DerivedF, typename DerivedS, typename DerivedI,
          typename DerivedC, typename DerivedN>
void signed_distance(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedV>& V,
                     const Eigen::MatrixBase<DerivedF>& F, const SignedDistanceType sign_type,
                     const typename DerivedV::Scalar lower_bound, const typename DerivedV::Scalar upper_bound,
                     Eigen::PlainObjectBase<DerivedS>& S, Eigen::PlainObjectBase<DerivedI>& I,
                     Eigen::PlainObjectBase<DerivedC>& C, Eigen::PlainObjectBase<DerivedN>& N) {
  gwn_detail::signed_distance(P.derived(), V.derived(), F.derived(), (int)sign_type, (double)lower_bound,
                              (double)upper_bound, S.derived(), I.derived(), C.derived(), &N.derived());
}

template <typename DerivedP, typename DerivedV, typename DerivedF, typename DerivedS, typename DerivedI,
          typename DerivedC, typename DerivedN>
void signed_distance(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedV>& V,
                     const Eigen::MatrixBase<DerivedF>& F, const SignedDistanceType sign_type,
                     Eigen::PlainObjectBase<DerivedS>& S, Eigen::PlainObjectBase<DerivedI>& I,
                     Eigen::PlainObjectBase<DerivedC>& C, Eigen::PlainObjectBase<DerivedN>& N) {
  typedef typename DerivedV::Scalar Scalar;  // signed_distance.cpp:436-439
  signed_distance(P, V, F, sign_type, std::numeric_limits<Scalar>::min(), std::numeric_limits<Scalar>::max(), S, I,
                  C, N);
}

namespace gwn_detail {
// The soup whose vertices equal V (as doubles): `tree` when it is a built
// FastWindingNumberBVH, else fwn_bvh when it was built from the same
// vertices, else a tree built here from V (owned by `tmp`).
template <typename Tree, typename MV, typename MF>
gwn_soup* closest_point_soup(const Tree& tree, const FastWindingNumberBVH& fwn_bvh, const MV& V, const MF& F,
                             FastWindingNumberBVH& tmp) {
  if constexpr (std::is_same<Tree, FastWindingNumberBVH>::value) {
    if (tree.handle) return tree.handle;
  }
  const auto v = rows3<double>(V, "V");
  if (fwn_bvh.handle && fwn_bvh.num_vertices == (int64_t)V.rows() && fwn_bvh.vertex_hash == hash_doubles(v))
    return fwn_bvh.handle;
  build(V, F, 0, tmp);
  return tmp.handle;
}

// signed_distance.cpp:665-711 for the points of P: unsigned distance and
// closest point from `closest`, sign 1 - 2|w| with w = fast_winding_number(
// fwn_bvh, 2, p.cast<float>()) as a float
template <typename MP, typename MS>
void signed_distance_fwn(gwn_soup* closest, const FastWindingNumberBVH& fwn_bvh, const MP& P, MS& S) {
  const auto p = rows3<double>(P, "P");
  const int64_t n = P.rows();
  std::vector<double> d((size_t)n), c(3 * (size_t)n), w((size_t)n);
  std::vector<int32_t> i((size_t)n);
  if (n > 0) {
    check(gwn_signed_distance(closest, p.data(), n, GWN_SIGN_UNSIGNED, std::numeric_limits<double>::min(),
                              std::numeric_limits<double>::max(), d.data(), i.data(), c.data()));
    check(gwn_soup_query(fwn_bvh.handle, p.data(), n, 2.0f, GWN_PREC_FP32, w.data()));
  }
  S.resize(n, 1);
  for (int64_t k = 0; k < n; ++k) {
    const double wf = (double)(float)w[k];
    S(k, 0) = static_cast<typename std::decay<decltype(S(0, 0))>::type>((1. - 2. * std::fabs(wf)) * d[k]);
  }
}
}  // namespace gwn_detail

// signed_distance.cpp:665-687: closest points from `tree` when it is a
// FastWindingNumberBVH over (V, F), else from fwn_bvh when that was built
// from V itself, else from a tree built here from V (the reference's AABB is
// over the caller's V; fwn_bvh may hold other, e.g. float-rounded, vertices);
// the sign from fwn_bvh at accuracy_scale 2; S = sqrt(d2) (1 - 2 |w|)
template <typename DerivedP, typename DerivedV, typename DerivedF, typename Tree, typename DerivedS>
void signed_distance_fast_winding_number(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedV>& V,
                                         const Eigen::MatrixBase<DerivedF>& F, const Tree& tree,
                                         const FastWindingNumberBVH& fwn_bvh, Eigen::PlainObjectBase<DerivedS>& S) {
  if (!fwn_bvh.handle) throw std::runtime_error("FastWindingNumberBVH not built");
  FastWindingNumberBVH tmp;
  gwn_soup* closest = gwn_detail::closest_point_soup(tree, fwn_bvh, V.derived(), F.derived(), tmp);
  gwn_detail::signed_distance_fwn(closest, fwn_bvh, P.derived(), S.derived());
}

// signed_distance.cpp:688-711: one point, the same closest-point tree choice
template <typename Derivedq, typename DerivedV, typename DerivedF, typename Tree>
typename DerivedV::Scalar signed_distance_fast_winding_number(const Eigen::MatrixBase<Derivedq>& q,
                                                              const Eigen::MatrixBase<DerivedV>& V,
                                                              const Eigen::MatrixBase<DerivedF>& F, const Tree& tree,
                                                              const FastWindingNumberBVH& fwn_bvh) {
  if (!fwn_bvh.handle) throw std::runtime_error("FastWindingNumberBVH not built");
  FastWindingNumberBVH tmp;
  gwn_soup* closest = gwn_detail::closest_point_soup(tree, fwn_bvh, V.derived(), F.derived(), tmp);
  Eigen::Matrix<double, Eigen::Dynamic, 1> S1;
  Eigen::Matrix<double, 1, 3> q1;
  for (int j = 0; j < 3; ++j) q1(0, j) = static_cast<double>(q.derived()(0, j));
  gwn_detail::signed_distance_fwn(closest, fwn_bvh, q1, S1);
  return static_cast<typename DerivedV::Scalar>(S1(0, 0));
}

// signed_distance.cpp:579-629 (3D): closest point and s = 1 - 2 w with the
// exact winding number.  `tree` and `hier` stand for the reference's AABB and
// WindingNumberAABB: a FastWindingNumberBVH built over (V, F) passed as `tree`
// serves the closest point (anything else, e.g. 0, makes the call build one);
// an igl::WindingNumberAABB passed as `hier` gives w (hier.winding_number(q),
// as the reference); anything else sums over every face.
template <typename Tree, typename DerivedV, typename DerivedF, typename Hier, typename Derivedq, typename Scalar,
          typename Derivedc>
void signed_distance_winding_number(const Tree& tree, const Eigen::MatrixBase<DerivedV>& V,
                                    const Eigen::MatrixBase<DerivedF>& F, const Hier& hier,
                                    const Eigen::MatrixBase<Derivedq>& q, Scalar& s, Scalar& sqrd, int& i,
                                    Eigen::PlainObjectBase<Derivedc>& c) {
  const gwn_exact_tree* h = nullptr;
  if constexpr (is_winding_number_aabb<Hier>::value) h = hier.tree();
  else (void)hier;
  double sd = 0.0, d2 = 0.0, cc[3];
  if constexpr (std::is_same<Tree, FastWindingNumberBVH>::value) {
    if (!tree.handle) throw std::runtime_error("FastWindingNumberBVH not built");
    gwn_detail::signed_distance_wn_point(tree.handle, h, V.derived(), F.derived(), q.derived(), sd, d2, i, cc);
  } else {
    FastWindingNumberBVH bvh;
    gwn_detail::build(V.derived(), F.derived(), 0, bvh);
    gwn_detail::signed_distance_wn_point(bvh.handle, h, V.derived(), F.derived(), q.derived(), sd, d2, i, cc);
  }
  s = static_cast<Scalar>(sd);
  sqrd = static_cast<Scalar>(d2);
  auto& cd = c.derived();
  cd.resize(1, 3);
  for (int j = 0; j < 3; ++j) cd(0, j) = static_cast<typename std::decay<decltype(cd(0, 0))>::type>(cc[j]);
}

template <typename Tree, typename DerivedV, typename DerivedF, typename Hier, typename Derivedq>
typename DerivedV::Scalar signed_distance_winding_number(const Tree& tree, const Eigen::MatrixBase<DerivedV>& V,
                                                         const Eigen::MatrixBase<DerivedF>& F, const Hier& hier,
                                                         const Eigen::MatrixBase<Derivedq>& q) {
  typedef typename DerivedV::Scalar Scalar;
  Scalar s = 0, sqrd = 0;
  int i = -1;
  Eigen::Matrix<double, Eigen::Dynamic, Eigen::Dynamic> c;
  signed_distance_winding_number(tree, V, F, hier, q, s, sqrd, i, c);
  return s * std::sqrt(sqrd);
}

template <typename DerivedV, typename DerivedF>
void fast_winding_number(const Eigen::MatrixBase<DerivedV>& V, const Eigen::MatrixBase<DerivedF>& F, const int order,
                         FastWindingNumberBVH& fwn_bvh) {
  gwn_detail::build(V.derived(), F.derived(), order, fwn_bvh);
}

template <typename DerivedQ, typename DerivedW>
void fast_winding_number(const FastWindingNumberBVH& fwn_bvh, const float accuracy_scale,
                         const Eigen::MatrixBase<DerivedQ>& Q, Eigen::PlainObjectBase<DerivedW>& W) {
  gwn_detail::query(fwn_bvh, accuracy_scale, Q.derived(), W.derived());
}

template <typename Derivedp>
typename Derivedp::Scalar fast_winding_number(const FastWindingNumberBVH& fwn_bvh, const float accuracy_scale,
                                              const Eigen::MatrixBase<Derivedp>& p) {
  return static_cast<typename Derivedp::Scalar>(gwn_detail::query_point(fwn_bvh, accuracy_scale, p.derived()));
}

template <typename DerivedV, typename DerivedF, typename DerivedQ, typename DerivedW>
void fast_winding_number(const Eigen::MatrixBase<DerivedV>& V, const Eigen::MatrixBase<DerivedF>& F,
                         const Eigen::MatrixBase<DerivedQ>& Q, Eigen::PlainObjectBase<DerivedW>& W) {
  FastWindingNumberBVH bvh;
  gwn_detail::build(V.derived(), F.derived(), 2, bvh);
  gwn_detail::query(bvh, 2.0f, Q.derived(), W.derived());
}

template <typename DerivedP, typename DerivedA, typename DerivedN, typename DerivedQ, typename BetaType,
          typename DerivedWN>
void fast_winding_number(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedN>& N,
                         const Eigen::MatrixBase<DerivedA>& A, const Eigen::MatrixBase<DerivedQ>& Q,
                         const int expansion_order, const BetaType beta, Eigen::PlainObjectBase<DerivedWN>& WN) {
  gwn_detail::cloud(P.derived(), N.derived(), A.derived(), Q.derived(), expansion_order, (double)beta, WN.derived());
}

template <typename DerivedP, typename DerivedA, typename DerivedN, typename DerivedQ, typename DerivedWN>
void fast_winding_number(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedN>& N,
                         const Eigen::MatrixBase<DerivedA>& A, const Eigen::MatrixBase<DerivedQ>& Q,
                         Eigen::PlainObjectBase<DerivedWN>& WN) {
  gwn_detail::cloud(P.derived(), N.derived(), A.derived(), Q.derived(), 2, 2.0, WN.derived());
}

template <typename DerivedP, typename IndexType, typename DerivedCH, typename DerivedCN, typename DerivedW,
          typename DerivedI>
void knn(const Eigen::MatrixBase<DerivedP>& P, size_t k, const std::vector<std::vector<IndexType>>& point_indices,
         const Eigen::MatrixBase<DerivedCH>& CH, const Eigen::MatrixBase<DerivedCN>& CN,
         const Eigen::MatrixBase<DerivedW>& W, Eigen::PlainObjectBase<DerivedI>& I) {
  (void)CN;
  (void)W;
  gwn_detail::knn(P.derived(), P.derived(), k, point_indices, CH.derived(), I.derived());
}

template <typename DerivedP, typename DerivedV, typename IndexType, typename DerivedCH, typename DerivedCN,
          typename DerivedW, typename DerivedI>
void knn(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedV>& V, size_t k,
         const std::vector<std::vector<IndexType>>& point_indices, const Eigen::MatrixBase<DerivedCH>& CH,
         const Eigen::MatrixBase<DerivedCN>& CN, const Eigen::MatrixBase<DerivedW>& W,
         Eigen::PlainObjectBase<DerivedI>& I) {
  (void)CN;  // the candidates and distances fix the answer; the cells' boxes only prune
  (void)W;
  gwn_detail::knn(P.derived(), V.derived(), k, point_indices, CH.derived(), I.derived());
}

// igl::WindingNumberAntipodalScene (WindingNumberAntipodalScene.h:48-256).
// The intersector arguments of winding_number are accepted and not used: the
// scene counts the ray crossings on the GPU over the mesh's own tree.  Like
// the reference's scene it keeps no reference to V and F.
template <typename Scalar>
class WindingNumberAntipodalScene {
 public:
  using Point = Eigen::Matrix<Scalar, 1, 3>;
  using Direction = Eigen::Matrix<Scalar, 1, 3>;

  template <typename DerivedV, typename DerivedF>
  WindingNumberAntipodalScene(const Eigen::MatrixBase<DerivedV>& V, const Eigen::MatrixBase<DerivedF>& F,
                              const Direction& x0 = default_x0())
      : m_x0(x0) {
    const double d[3] = {(double)x0(0, 0), (double)x0(0, 1), (double)x0(0, 2)};
    m_scene = gwn_detail::antipodal_build(V.derived(), F.derived(), d);
    int64_t nb = 0, nf = 0;
    gwn_detail::check(gwn_antipodal_info(m_scene, &nb, &nf, nullptr));
    m_nb = (size_t)nb;
    m_nf = (size_t)nf;
  }
  WindingNumberAntipodalScene(const WindingNumberAntipodalScene&) = delete;
  WindingNumberAntipodalScene& operator=(const WindingNumberAntipodalScene&) = delete;
  ~WindingNumberAntipodalScene() {
    if (m_scene) gwn_antipodal_free(m_scene);
  }

  template <typename Intersector, typename Derivedp>
  Scalar winding_number(const Intersector& intersector, const Eigen::MatrixBase<Derivedp>& p) const {
    (void)intersector;
    Eigen::Matrix<double, Eigen::Dynamic, Eigen::Dynamic> O(1, 3);
    for (int j = 0; j < 3; ++j) O(0, j) = static_cast<double>(p.derived()(0, j));
    Eigen::Matrix<double, Eigen::Dynamic, 1> W;
    gwn_detail::antipodal_query(m_scene, O, W);
    return static_cast<Scalar>(W(0, 0));
  }

  template <typename Intersector, typename DerivedO, typename DerivedW>
  void winding_number(const Intersector& intersector, const Eigen::MatrixBase<DerivedO>& O,
                      Eigen::PlainObjectBase<DerivedW>& W) const {
    (void)intersector;
    gwn_detail::antipodal_query(m_scene, O.derived(), W.derived());
  }

  const Direction& x0() const { return m_x0; }
  size_t num_boundary_segments() const { return m_nb; }
  size_t num_faces() const { return m_nf; }

  // normalize(1, sqrt 2, sqrt 3) (WindingNumberAntipodalScene.h:167-173)
  static Direction default_x0() {
    const double n = std::sqrt((1.0 + 2.0) + 3.0);
    Direction d;
    d(0, 0) = static_cast<Scalar>(1.0 / n);
    d(0, 1) = static_cast<Scalar>(std::sqrt(2.0) / n);
    d(0, 2) = static_cast<Scalar>(std::sqrt(3.0) / n);
    return d;
  }

 private:
  gwn_antipodal* m_scene = nullptr;
  Direction m_x0;
  size_t m_nb = 0, m_nf = 0;
};

// winding_number_antipodal.h:43-53
template <typename Intersector, typename DerivedV, typename DerivedF, typename DerivedO, typename DerivedW>
void winding_number_antipodal(const Eigen::MatrixBase<DerivedV>& V, const Eigen::MatrixBase<DerivedF>& F,
                              const Intersector& intersector, const Eigen::MatrixBase<DerivedO>& O,
                              Eigen::PlainObjectBase<DerivedW>& W) {
  (void)intersector;
  gwn_antipodal* a = gwn_detail::antipodal_build(V.derived(), F.derived(), nullptr);
  try {
    gwn_detail::antipodal_query(a, O.derived(), W.derived());
  } catch (...) {
    gwn_antipodal_free(a);
    throw;
  }
  gwn_antipodal_free(a);
}

namespace copyleft {
namespace cgal {

template <typename DerivedP, typename DerivedI, typename DerivedN, typename DerivedA, typename DerivedT>
void point_areas(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedI>& I,
                 const Eigen::MatrixBase<DerivedN>& N, Eigen::PlainObjectBase<DerivedA>& A,
                 Eigen::PlainObjectBase<DerivedT>& T) {
  gwn_detail::point_areas(P.derived(), I.derived(), N.derived(), A.derived(), T.derived());
}

template <typename DerivedP, typename DerivedI, typename DerivedN, typename DerivedA>
void point_areas(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedI>& I,
                 const Eigen::MatrixBase<DerivedN>& N, Eigen::PlainObjectBase<DerivedA>& A) {
  Eigen::MatrixXd T;
  gwn_detail::point_areas(P.derived(), I.derived(), N.derived(), A.derived(), T);
}

template <typename DerivedP, typename DerivedN, typename DerivedQ, typename BetaType, typename DerivedWN>
void fast_winding_number(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedN>& N,
                         const Eigen::MatrixBase<DerivedQ>& Q, const int expansion_order, const BetaType beta,
                         Eigen::PlainObjectBase<DerivedWN>& WN) {
  gwn_detail::cloud_estimated_areas(P.derived(), N.derived(), Q.derived(), expansion_order, (double)beta,
                                    WN.derived());
}

template <typename DerivedP, typename DerivedN, typename DerivedQ, typename DerivedWN>
void fast_winding_number(const Eigen::MatrixBase<DerivedP>& P, const Eigen::MatrixBase<DerivedN>& N,
                         const Eigen::MatrixBase<DerivedQ>& Q, Eigen::PlainObjectBase<DerivedWN>& WN) {
  gwn_detail::cloud_estimated_areas(P.derived(), N.derived(), Q.derived(), 2, 2.0, WN.derived());
}

}  // namespace cgal
}  // namespace copyleft

}  // namespace igl
#endif
#endif
