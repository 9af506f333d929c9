// gwn_igl.h -- libigl-signature facade over the gwn C ABI (header only).
//
// Replaces, with the same names, argument meaning and output resizing:
//   igl::winding_number(V,F,O,W)                      winding_number.h:41-48
//   igl::fast_winding_number(V,F,Q,W)                 fast_winding_number.h:160-165
//   igl::fast_winding_number(V,F,order,bvh)           fast_winding_number.h:176-180
//   igl::fast_winding_number(bvh,accuracy_scale,Q,W)  fast_winding_number.h:191-196
//   igl::fast_winding_number(bvh,accuracy_scale,p)    fast_winding_number.h:203-206
//   igl::fast_winding_number(P,N,A,Q,order,beta,WN)   fast_winding_number.h:111-118
//   igl::fast_winding_number(P,N,A,Q,WN)              fast_winding_number.h:126-131
//   struct igl::FastWindingNumberBVH                  fast_winding_number.h:140-152
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
#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "gwn.h"

namespace igl {

// igl::FastWindingNumberBVH: owns the GPU tree.  Like the reference's (which
// holds raw pointers into itself) it is neither copyable nor movable.
struct FastWindingNumberBVH {
  gwn_soup* handle = nullptr;
  int tree_kind = GWN_TREE_SAH4;   // reference topology by default
  int precision = GWN_PREC_FP32;   // reference arithmetic by default
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

template <typename MV, typename MF, typename MO, typename MW>
void winding_number(const MV& V, const MF& F, const MO& O, MW& W) {
  const auto v = rows3<double>(V, "V");
  const auto f = faces(F);
  const auto o = rows3<double>(O, "O");
  std::vector<double> w((size_t)O.rows());
  check(gwn_exact_winding_number(v.data(), V.rows(), f.data(), F.rows(), o.data(), O.rows(), w.data()));
  W.resize(O.rows(), 1);
  for (int64_t i = 0; i < (int64_t)O.rows(); ++i) W(i, 0) = static_cast<typename std::decay<decltype(W(0, 0))>::type>(w[i]);
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

}  // namespace gwn_detail
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

}  // namespace igl
#endif
#endif
