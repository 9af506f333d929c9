// Exercises include/gwn_igl.h through its generic conversion layer with a
// minimal column-major matrix type (the Eigen overloads forward to the same
// templates).  Built by tests/test_facade.py; run on the GPU.
#include <cmath>
#include <cstdio>
#include <limits>
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
  // winding_number(V,F,p): single point (winding_number.h:49-60)
  Mat<double> pin(1, 3);
  pin(0, 0) = 0.2, pin(0, 1) = 0.2, pin(0, 2) = 0.2;
  if (std::fabs(igl::gwn_detail::winding_number_point(V, F, pin) - 1.0) > 1e-15) bad |= 64;
  // octree -> precompute -> eval over the caller's octree reproduces the
  // one-shot overload: octree.h:36-41, fast_winding_number.h:44-53 / 88-99
  {
    Mat<double> Pc(9, 3), Nc(9, 3), Ac(9, 1), Qq(3, 3);
    for (int i = 0; i < 8; ++i) {
      Pc(i, 0) = (i & 1) ? 1 : -1, Pc(i, 1) = (i & 2) ? 1 : -1, Pc(i, 2) = (i & 4) ? 1 : -1;
      for (int j = 0; j < 3; ++j) Nc(i, j) = Pc(i, j) / std::sqrt(3.0);
      Ac(i, 0) = 0.5;
    }
    Pc(8, 0) = 0.9, Pc(8, 1) = 0.8, Pc(8, 2) = 0.85;
    Nc(8, 0) = 1, Nc(8, 1) = 0, Nc(8, 2) = 0;
    Ac(8, 0) = 0.25;
    const double qq[3][3] = {{0, 0, 0}, {3, 0.5, -1}, {0.95, 0.9, 0.9}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Qq(i, j) = qq[i][j];
    std::vector<std::vector<int>> PI;
    Mat<int> CH;
    Mat<double> CN, Wc;
    igl::gwn_detail::octree(Pc, PI, CH, CN, Wc);
    // root: all 9 points, split into 8 children; the octant of (1,1,1) holds 2
    if (CH.rows() < 9 || CH(0, 0) != 1 || PI[0].size() != 9 || PI[8].size() != 2 || std::fabs(Wc(0, 0) - 2) > 0)
      bad |= 128;
    Mat<double> CM, R, EC, W1, W2;
    igl::gwn_detail::cloud_precompute(Pc, Nc, Ac, PI, CH, 2, CM, R, EC);
    if (CM.rows() != CH.rows() || EC.cols() != 39 || R.rows() != CH.rows()) bad |= 256;
    igl::gwn_detail::cloud_eval(Pc, Nc, Ac, PI, CH, CM, R, EC, Qq, 2.0, W1);
    igl::gwn_detail::cloud(Pc, Nc, Ac, Qq, 2, 2.0, W2);
    for (int i = 0; i < 3; ++i)
      if (std::fabs(W1(i, 0) - W2(i, 0)) > 1e-6) bad |= 512;
  }
  // signed_distance (signed_distance.cpp:4-44 single tet): S = sqrt(1/12)
  {
    Mat<double> Pt(1, 3), S, C;
    Mat<int> I;
    Pt(0, 0) = 0.5, Pt(0, 1) = 0.5, Pt(0, 2) = 0.5;
    igl::gwn_detail::signed_distance(Pt, V, F, igl::SIGNED_DISTANCE_TYPE_FAST_WINDING_NUMBER,
                                     std::numeric_limits<double>::min(), std::numeric_limits<double>::max(), S, I, C);
    if (S.rows() != 1 || std::fabs(S(0, 0) - std::sqrt(1.0 / 12.0)) > 1e-7 || I(0, 0) != 3) bad |= 1024;
  }
  // UT_SubtendedAngle (2D): a closed 64-gon, counter-clockwise: 2 pi inside,
  // 0 outside; the batch form agrees with the per-point one
  {
    using igl::FastWindingNumber::HDK_Sample::UT_SubtendedAngle;
    using igl::FastWindingNumber::HDK_Sample::UT_Vector2T;
    const int n = 64;
    std::vector<UT_Vector2T<float>> P(n);
    std::vector<int> E(2 * n);
    for (int i = 0; i < n; ++i) {
      const double t = 2 * 3.14159265358979323846 * i / n;
      P[i] = UT_Vector2T<float>((float)std::cos(t), (float)std::sin(t));
      E[2 * i] = i;
      E[2 * i + 1] = (i + 1) % n;
    }
    UT_SubtendedAngle<float, float> sa(n, E.data(), n, P.data(), 2);
    const float in = sa.computeAngle(UT_Vector2T<float>(0.1f, -0.2f), 2.0f);
    const float out = sa.computeAngle(UT_Vector2T<float>(3.0f, 0.5f), 2.0f);
    if (std::fabs(in / (2 * 3.14159265f) - 1.0f) > 2e-3f || std::fabs(out) > 2e-2f) bad |= 2048;
    const UT_Vector2T<float> qs[2] = {UT_Vector2T<float>(0.1f, -0.2f), UT_Vector2T<float>(3.0f, 0.5f)};
    float both[2];
    sa.computeAngles(2, qs, both, 2.0f);
    if (both[0] != in || both[1] != out || sa.isClear()) bad |= 2048;
    UT_SubtendedAngle<float, float> empty;
    if (!empty.isClear() || empty.computeAngle(UT_Vector2T<float>(0.0f, 0.0f)) != 0.0f) bad |= 2048;
  }
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
