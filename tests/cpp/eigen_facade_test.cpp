// Calls every Eigen-signature overload of include/gwn_igl.h (compiled against
// the mock Eigen/Core in tests/cpp/mock_eigen) with the reference's own test
// inputs: tests/include/igl/signed_distance.cpp:4-44 (single tet) and
// tests/include/igl/fast_winding_number.cpp:17-53 (one-point cloud).
#include <cmath>
#include <cstdio>
#include <vector>

#include "gwn_igl.h"

int main() {
  int bad = 0;
  Eigen::MatrixXd V(4, 3);
  const double v[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) V(i, j) = v[i][j];
  Eigen::MatrixXi F(4, 3);
  const int f[4][3] = {{0, 1, 3}, {0, 2, 1}, {0, 3, 2}, {1, 2, 3}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) F(i, j) = f[i][j];
  Eigen::MatrixXd P(1, 3);
  P(0, 0) = 0.5, P(0, 1) = 0.5, P(0, 2) = 0.5;
  const double Sexact = std::sqrt(1.0 / 12.0);
  // igl::signed_distance, both overloads, every provided sign type
  for (igl::SignedDistanceType t : {igl::SIGNED_DISTANCE_TYPE_PSEUDONORMAL, igl::SIGNED_DISTANCE_TYPE_WINDING_NUMBER,
                                    igl::SIGNED_DISTANCE_TYPE_DEFAULT, igl::SIGNED_DISTANCE_TYPE_UNSIGNED,
                                    igl::SIGNED_DISTANCE_TYPE_FAST_WINDING_NUMBER}) {
    Eigen::VectorXd S;
    Eigen::VectorXi I;
    Eigen::MatrixXd C, N;
    igl::signed_distance(P, V, F, t, S, I, C, N);
    const double tol = t == igl::SIGNED_DISTANCE_TYPE_FAST_WINDING_NUMBER ? 1e-7 : 1e-15;
    if (S.rows() != 1 || std::fabs(S(0, 0) - Sexact) > tol || I(0, 0) != 3 || C.cols() != 3) bad |= 1;
    // the pseudonormal type resizes and fills N: the outward normal of face 3
    if (t == igl::SIGNED_DISTANCE_TYPE_PSEUDONORMAL &&
        (N.rows() != 1 || N.cols() != 3 || std::fabs(N(0, 0) - std::sqrt(1.0 / 3.0)) > 1e-15 ||
         std::fabs(N(0, 2) - std::sqrt(1.0 / 3.0)) > 1e-15))
      bad |= 1 << 23;
    // bounds (0.05, 0.1): squared distances kept in [0.05^2, 0.1^2) (signed_distance.cpp:79-84);
    // d^2 = 1/12 is outside: NaN, I = F.rows() + 1
    igl::signed_distance(P, V, F, t, 0.05, 0.1, S, I, C, N);
    if (!std::isnan(S(0, 0)) || I(0, 0) != 5) bad |= 2;
  }
  // signed_distance_fast_winding_number with a prebuilt bvh
  {
    igl::FastWindingNumberBVH bvh;
    igl::fast_winding_number(V, F, 2, bvh);
    Eigen::VectorXd S;
    const int tree = 0;  // the caller's AABB is not used
    igl::signed_distance_fast_winding_number(P, V, F, tree, bvh, S);
    if (S.rows() != 1 || std::fabs(S(0, 0) - Sexact) > 1e-7) bad |= 4;
    // the single-point form (signed_distance.cpp:688-711)
    Eigen::Matrix<double, 1, 3> q1;
    q1(0, 0) = 0.5, q1(0, 1) = 0.5, q1(0, 2) = 0.5;
    const double sq1 = igl::signed_distance_fast_winding_number(q1, V, F, tree, bvh);
    if (std::fabs(sq1 - Sexact) > 1e-7) {
      std::printf("signed_distance_fast_winding_number(q): %.17g\n", sq1);
      bad |= 1 << 17;
    }
    // signed_distance_winding_number (signed_distance.cpp:579-629), with the
    // bvh as the tree and with a tree built by the call
    double s1 = 0, sqrd1 = 0;
    int i1 = -1;
    Eigen::MatrixXd c1;
    igl::signed_distance_winding_number(bvh, V, F, 0, q1, s1, sqrd1, i1, c1);
    // outside, w is zero up to the rounding of the solid angles: s = 1 - 2w
    if (std::fabs(s1 - 1.0) > 1e-14 || std::fabs(sqrd1 - 1.0 / 12.0) > 1e-15 || i1 != 3 ||
        std::fabs(c1(0, 0) - 1.0 / 3.0) > 1e-15) {
      std::printf("signed_distance_winding_number: s %.17g sqrd %.17g i %d c %.17g\n", s1, sqrd1, i1, c1(0, 0));
      bad |= 1 << 18;
    }
    Eigen::Matrix<double, 1, 3> q2;
    q2(0, 0) = 0.1, q2(0, 1) = 0.1, q2(0, 2) = 0.2;  // inside: w = 1, s = -1
    const double sq2 = igl::signed_distance_winding_number(0, V, F, 0, q2);
    if (std::fabs(sq2 + 0.1) > 1e-14) {
      std::printf("signed_distance_winding_number(q): %.17g\n", sq2);
      bad |= 1 << 19;
    }
    // the batch and single-point FWN overloads
    Eigen::MatrixXd Q(2, 3);
    Q(0, 0) = 0.1, Q(0, 1) = 0.1, Q(0, 2) = 0.1;
    Q(1, 0) = 2.0, Q(1, 1) = 0.0, Q(1, 2) = 0.0;
    Eigen::VectorXf W;
    igl::fast_winding_number(bvh, 2.0f, Q, W);
    if (W.rows() != 2 || std::fabs(W(0, 0) - 1.0f) > 1e-6f || std::fabs(W(1, 0)) > 1e-3f) bad |= 8;
    Eigen::Matrix<double, 1, 3> p;
    p(0, 0) = 0.1, p(0, 1) = 0.1, p(0, 2) = 0.1;
    if (std::fabs(igl::fast_winding_number(bvh, 2.0f, p) - 1.0) > 1e-6) bad |= 16;
  }
  // exact winding number, batch and single point
  {
    Eigen::VectorXd W;
    igl::winding_number(V, F, P, W);
    Eigen::Matrix<double, 1, 3> p;
    p(0, 0) = 0.1, p(0, 1) = 0.1, p(0, 2) = 0.1;
    if (std::fabs(W(0, 0)) > 1e-15 || std::fabs(igl::winding_number(V, F, p) - 1.0) > 1e-15) bad |= 32;
    Eigen::VectorXd W2;
    igl::fast_winding_number(V, F, P, W2);
    if (std::fabs(W2(0, 0)) > 1e-6) bad |= 64;
  }
  // point cloud: one-shot, default, octree + precompute + eval
  {
    Eigen::MatrixXd Pc(1, 3), Nc(1, 3), Ac(1, 1), Qc(4, 3);
    Pc(0, 0) = 0.1, Pc(0, 1) = 0.2, Pc(0, 2) = 0.3;
    const double nn = std::sqrt(16.0 + 25.0 + 36.0);
    Nc(0, 0) = 4 / nn, Nc(0, 1) = 5 / nn, Nc(0, 2) = 6 / nn;
    Ac(0, 0) = 0.7;
    const double qc[4][3] = {{0, 0, 0}, {4, -3, 2}, {-1, 1, 3}, {0, 5, -2}};
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j) Qc(i, j) = qc[i][j];
    const double gold[4] = {0.38779369004261133, -0.00041235296362485, -0.00362978253577090, -0.00041235296362485};
    Eigen::VectorXd W1, W2, W3;
    igl::fast_winding_number(Pc, Nc, Ac, Qc, 2, 2.0, W1);
    igl::fast_winding_number(Pc, Nc, Ac, Qc, W2);
    std::vector<std::vector<int>> PI;
    Eigen::MatrixXi CH;
    Eigen::MatrixXd CN, Wc, CM, R, EC;
    igl::octree(Pc, PI, CH, CN, Wc);
    igl::fast_winding_number(Pc, Nc, Ac, PI, CH, 2, CM, R, EC);
    igl::fast_winding_number(Pc, Nc, Ac, PI, CH, CM, R, EC, Qc, 2.0, W3);
    for (int i = 0; i < 4; ++i)
      if (std::fabs(W1(i, 0) - gold[i]) > 4e-7 || std::fabs(W2(i, 0) - gold[i]) > 4e-7 ||
          std::fabs(W3(i, 0) - gold[i]) > 4e-7)
        bad |= 128;
    if (CH.rows() != 1 || PI.size() != 1 || EC.cols() != 39) bad |= 256;
  }
  // read_triangle_mesh: a quad written here reads as two triangles
  {
    const char* path = "/tmp/gwn_eigen_facade_quad.obj";
    FILE* fp = std::fopen(path, "w");
    if (fp) {
      std::fputs("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n", fp);
      std::fclose(fp);
    }
    Eigen::MatrixXd Vm;
    Eigen::MatrixXi Fm;
    if (!igl::read_triangle_mesh(path, Vm, Fm) || Vm.rows() != 4 || Fm.rows() != 2 || Fm(1, 2) != 3) bad |= 512;
    std::remove(path);
    if (igl::read_triangle_mesh("/tmp/does_not_exist.xyz", Vm, Fm)) bad |= 1024;
  }
  // 2D winding number of a counter-clockwise square (F.cols() == 2)
  {
    Eigen::MatrixXd V2(4, 2), O2(2, 2);
    Eigen::MatrixXi E2(4, 2);
    const double sq[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int i = 0; i < 4; ++i) {
      V2(i, 0) = sq[i][0], V2(i, 1) = sq[i][1];
      E2(i, 0) = i, E2(i, 1) = (i + 1) % 4;
    }
    O2(0, 0) = 0.5, O2(0, 1) = 0.5, O2(1, 0) = 2.0, O2(1, 1) = 0.5;
    Eigen::VectorXd W2;
    igl::winding_number(V2, E2, O2, W2);
    if (W2.rows() != 2 || std::fabs(W2(0, 0) - 1.0) > 1e-15 || std::fabs(W2(1, 0)) > 1e-15) bad |= 2048;
  }
  // knn over an octree: tests/include/igl/knn.cpp:24-49 (cube corners, with a
  // tie resolved to the lower index)
  {
    Eigen::MatrixXd Vk(8, 3), Pk(3, 3);
    const double vk[8][3] = {{0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}};
    const double pk[3][3] = {{0, 0, 0.6}, {0.3, 0.1, 0.2}, {0.7, 0.6, 0}};
    const int ik[3][3] = {{0, 3, 2}, {3, 4, 0}, {5, 4, 1}};
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 3; ++j) Vk(i, j) = vk[i][j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Pk(i, j) = pk[i][j];
    std::vector<std::vector<int>> PI;
    Eigen::MatrixXi CH, I;
    Eigen::MatrixXd CN, Wk;
    igl::octree(Vk, PI, CH, CN, Wk);
    igl::knn(Pk, Vk, 3, PI, CH, CN, Wk, I);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (I.rows() != 3 || I.cols() != 3 || I(i, j) != ik[i][j]) bad |= 4096;
    igl::knn(Vk, 2, PI, CH, CN, Wk, I);  // V = P: each point is its own nearest
    for (int i = 0; i < 8; ++i)
      if (I(i, 0) != i) bad |= 4096;
  }
  // point areas on a flat triangular lattice: the centre's Voronoi area is
  // (sqrt 3 / 2) a^2; then the cloud FWN with estimated areas
  {
    const double a = 0.5;
    std::vector<double> pts;
    for (int i = -4; i <= 4; ++i)
      for (int j = -4; j <= 4; ++j) {
        pts.push_back(a * (i + 0.5 * j));
        pts.push_back(a * j * std::sqrt(3.0) / 2);
        pts.push_back(0.25);
      }
    const int n = (int)pts.size() / 3;
    Eigen::MatrixXd Pl(n, 3), Nl(n, 3), Al, Tl;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < 3; ++j) {
        Pl(i, j) = pts[3 * i + j];
        Nl(i, j) = j == 2 ? 1.0 : 0.0;
      }
    std::vector<std::vector<int>> PI;
    Eigen::MatrixXi CH, I;
    Eigen::MatrixXd CN, Wl;
    igl::octree(Pl, PI, CH, CN, Wl);
    igl::knn(Pl, 21, PI, CH, CN, Wl, I);
    igl::copyleft::cgal::point_areas(Pl, I, Nl, Al, Tl);
    const int c = n / 2;  // (0, 0)
    if (std::fabs(Al(c, 0) - std::sqrt(3.0) / 2 * a * a) > 1e-12 || std::fabs(Tl(c, 2) - 1.0) > 1e-12) bad |= 8192;
    Eigen::MatrixXd A1;
    igl::copyleft::cgal::point_areas(Pl, I, Nl, A1);
    if (A1(c, 0) != Al(c, 0)) bad |= 8192;
    // a closed cloud: 12 icosahedron vertices with outward normals
    const double t = (1 + std::sqrt(5.0)) / 2;
    const double ico[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                               {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    Eigen::MatrixXd Pi(12, 3), Ni(12, 3), Qi(2, 3);
    for (int i = 0; i < 12; ++i) {
      const double r = std::sqrt(1 + t * t);
      for (int j = 0; j < 3; ++j) Pi(i, j) = ico[i][j] / r, Ni(i, j) = ico[i][j] / r;
    }
    Qi(0, 0) = Qi(0, 1) = Qi(0, 2) = 0.0;
    Qi(1, 0) = 5.0, Qi(1, 1) = Qi(1, 2) = 0.0;
    Eigen::VectorXd W1, W2;
    igl::copyleft::cgal::fast_winding_number(Pi, Ni, Qi, 2, 2.0, W1);
    igl::copyleft::cgal::fast_winding_number(Pi, Ni, Qi, W2);
    if (W1.rows() != 2 || !(W1(0, 0) > 0.5) || !(std::fabs(W1(1, 0)) < 0.1) || W2(0, 0) != W1(0, 0)) bad |= 16384;
  }
  // antipodal method on the single tet (closed: an integer count) and on one
  // open triangle (tests/include/igl/embree/winding_number_antipodal.cpp:48-77)
  {
    struct NoIntersector {} none;
    Eigen::MatrixXd Pa(2, 3);
    Pa(0, 0) = 0.1, Pa(0, 1) = 0.1, Pa(0, 2) = 0.1;  // inside the tet
    Pa(1, 0) = 2.0, Pa(1, 1) = 0.3, Pa(1, 2) = 0.2;  // outside
    Eigen::VectorXd Wa;
    igl::winding_number_antipodal(V, F, none, Pa, Wa);
    if (Wa.rows() != 2 || Wa(0, 0) != 1.0 || Wa(1, 0) != 0.0) bad |= 32768;
    igl::WindingNumberAntipodalScene<double> scene(V, F);
    if (scene.num_boundary_segments() != 0 || scene.num_faces() != 4) bad |= 32768;
    Eigen::Matrix<double, 1, 3> p;
    p(0, 0) = 0.2, p(0, 1) = 0.2, p(0, 2) = 0.2;
    if (scene.winding_number(none, p) != 1.0) bad |= 32768;
    Eigen::MatrixXd Vt(3, 3);
    Eigen::MatrixXi Ft(1, 3);
    const double vt[3][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Vt(i, j) = vt[i][j];
    Ft(0, 0) = 0, Ft(0, 1) = 1, Ft(0, 2) = 2;
    Eigen::MatrixXd Pt(2, 3);
    Pt(0, 0) = Pt(0, 1) = 1.0 / 3.0, Pt(0, 2) = 0.5;
    Pt(1, 0) = Pt(1, 1) = 1.0 / 3.0, Pt(1, 2) = -0.5;
    igl::WindingNumberAntipodalScene<double> tri(Vt, Ft);
    Eigen::VectorXd Wt, Wx;
    tri.winding_number(none, Pt, Wt);
    igl::winding_number(Vt, Ft, Pt, Wx);
    if (tri.num_boundary_segments() != 3 || !(Wt(0, 0) * Wt(1, 0) < 0.0) || std::fabs(Wt(0, 0) - Wx(0, 0)) > 1e-12 ||
        std::fabs(Wt(1, 0) - Wx(1, 0)) > 1e-12)
      bad |= 65536;
  }
  // a bvh built from float-rounded vertices (as igl::fast_winding_number(
  // V.cast<float>(), F, 2, bvh) makes it): signed_distance_fast_winding_number
  // takes its closest points from the caller's double V, not the bvh's floats
  {
    Eigen::MatrixXd Vp(4, 3);
    const double vp[4][3] = {{0.1, 0.2, 0.3}, {1.1, 0.2, 0.3}, {0.1, 1.3, 0.3}, {0.1, 0.2, 1.7}};
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j) Vp(i, j) = vp[i][j];
    Eigen::Matrix<float, Eigen::Dynamic, Eigen::Dynamic> Vf(4, 3);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j) Vf(i, j) = (float)vp[i][j];
    igl::FastWindingNumberBVH bvh_f;
    igl::fast_winding_number(Vf, F, 2, bvh_f);
    Eigen::MatrixXd Pp(3, 3);
    const double pp[3][3] = {{0.9, 0.9, 0.9}, {0.2, 0.3, 0.35}, {-0.4, 0.5, 0.6}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Pp(i, j) = pp[i][j];
    Eigen::VectorXd Sf, Sd;
    Eigen::VectorXi Id;
    Eigen::MatrixXd Cd, Nd;
    igl::signed_distance_fast_winding_number(Pp, Vp, F, 0, bvh_f, Sf);
    igl::signed_distance(Pp, Vp, F, igl::SIGNED_DISTANCE_TYPE_FAST_WINDING_NUMBER, Sd, Id, Cd, Nd);
    for (int k = 0; k < 3; ++k)
      if (Sf.rows() != 3 || std::fabs(std::fabs(Sf(k, 0)) - std::fabs(Sd(k, 0))) > 1e-15 ||
          (Sf(k, 0) < 0) != (Sd(k, 0) < 0)) {
        std::printf("float-built bvh, point %d: %.17g vs %.17g\n", k, Sf(k, 0), Sd(k, 0));
        bad |= 1 << 20;
      }
  }
  // igl::WindingNumberAABB as `hier` of signed_distance_winding_number, and on its own
  {
    igl::WindingNumberAABB<double, int> hier(V, F);
    hier.grow();
    Eigen::Matrix<double, 1, 3> qi;
    qi(0, 0) = 0.1, qi(0, 1) = 0.1, qi(0, 2) = 0.2;
    if (std::fabs(hier.winding_number(qi) - 1.0) > 1e-14) bad |= 1 << 21;
    igl::FastWindingNumberBVH bvh;
    igl::fast_winding_number(V, F, 2, bvh);
    double s = 0, sqrd = 0;
    int i = -1;
    Eigen::MatrixXd c;
    igl::signed_distance_winding_number(bvh, V, F, hier, qi, s, sqrd, i, c);
    if (std::fabs(s + 1.0) > 1e-14 || std::fabs(std::sqrt(sqrd) - 0.1) > 1e-15) {
      std::printf("signed_distance_winding_number with hier: s %.17g sqrd %.17g\n", s, sqrd);
      bad |= 1 << 22;
    }
  }
  std::printf("eigen_facade_test: %s (flags %d)\n", bad ? "FAIL" : "OK", bad);
  return bad;
}
