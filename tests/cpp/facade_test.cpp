// Exercises include/gwn_igl.h through its generic conversion layer with a
// minimal column-major matrix type (the Eigen overloads forward to the same
// templates).  Built by tests/test_facade.py; run on the GPU.
#include <cmath>
#include <cstdio>
#include <vector>

#include "gwn_igl.h"

template <typename T>
struct Mat {  // column-major, like Eigen's default storage
  int64_t r = 0, c = 0;
  std::vector<T> d;
  Mat() = default;
  Mat(int64_t r_, int64_t c_) : r(r_), c(c_), d((size_t)(r_ * c_)) {}
  int64_t rows() const { return r; }
  int64_t cols() const { return c; }
  T& operator()(int64_t i, int64_t j) { return d[(size_t)(i + r * j)]; }
  const T& operator()(int64_t i, int64_t j) const { return d[(size_t)(i + r * j)]; }
  void resize(int64_t r_, int64_t c_) {
    r = r_;
    c = c_;
    d.assign((size_t)(r_ * c_), T(0));
  }
};

int main() {
  // tests/include/igl/signed_distance.cpp:4-44 tetrahedron
  Mat<double> V(4, 3);
  const double v[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) V(i, j) = v[i][j];
  Mat<int> F(4, 3);
  const int f[4][3] = {{0, 1, 3}, {0, 2, 1}, {0, 3, 2}, {1, 2, 3}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) F(i, j) = f[i][j];
  Mat<double> Q(3, 3);
  const double q[3][3] = {{0.5, 0.5, 0.5}, {0.1, 0.1, 0.1}, {0.2, 0.3, 0.1}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Q(i, j) = q[i][j];
  int bad = 0;
  Mat<double> W;
  igl::gwn_detail::winding_number(V, F, Q, W);
  if (W.rows() != 3 || W.cols() != 1) bad |= 1;
  if (std::fabs(W(0, 0)) > 1e-15 || std::fabs(W(1, 0) - 1) > 1e-15 || std::fabs(W(2, 0) - 1) > 1e-15) bad |= 2;
  igl::FastWindingNumberBVH bvh;
  igl::gwn_detail::build(V, F, 2, bvh);
  Mat<float> Wf;
  igl::gwn_detail::query(bvh, 2.0f, Q, Wf);
  // signed_distance.cpp:121-122: FWN sign at (0.5,0.5,0.5) within 1e-7
  if (Wf.rows() != 3 || std::fabs(Wf(0, 0)) > 1e-7 || std::fabs(Wf(1, 0) - 1) > 1e-6) bad |= 4;
  Mat<double> p(1, 3);
  p(0, 0) = 0.1;
  p(0, 1) = 0.1;
  p(0, 2) = 0.1;
  const double w1 = igl::gwn_detail::query_point(bvh, 2.0f, p);
  if (std::fabs(w1 - 1.0) > 1e-6) bad |= 8;
  // one-point cloud (tests/include/igl/fast_winding_number.cpp:17-53)
  Mat<double> P(1, 3), N(1, 3), A(1, 1), Qc(4, 3);
  P(0, 0) = 0.1, P(0, 1) = 0.2, P(0, 2) = 0.3;
  const double nn = std::sqrt(16.0 + 25.0 + 36.0);
  N(0, 0) = 4 / nn, N(0, 1) = 5 / nn, N(0, 2) = 6 / nn;
  A(0, 0) = 0.7;
  const double qc[4][3] = {{0, 0, 0}, {4, -3, 2}, {-1, 1, 3}, {0, 5, -2}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) Qc(i, j) = qc[i][j];
  Mat<double> WN;
  igl::gwn_detail::cloud(P, N, A, Qc, 2, 2.0, WN);
  const double gold[4] = {0.38779369004261133, -0.00041235296362485, -0.00362978253577090, -0.00041235296362485};
  for (int i = 0; i < 4; ++i)
    if (std::fabs(WN(i, 0) - gold[i]) > 1e-6 * 0.4) bad |= 16;
  // bad input is reported, not asserted
  Mat<int> Fbad(1, 3);
  Fbad(0, 0) = 0, Fbad(0, 1) = 1, Fbad(0, 2) = 9;
  try {
    igl::gwn_detail::build(V, Fbad, 2, bvh);
    bad |= 32;
  } catch (const std::runtime_error&) {
  }
  std::printf("facade_test: %s (flags %d) W=%.3g %.9g fwn=%.3g %.9g cloud=%.9g\n", bad ? "FAIL" : "OK", bad, W(0, 0),
              W(1, 0), (double)Wf(0, 0), (double)Wf(1, 0), WN(0, 0));
  return bad;
}
