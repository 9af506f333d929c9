// normals.cpp -- the normals of the pseudonormal sign test, built on the host.
//
// signed_distance.cpp:68-76 computes, for SIGNED_DISTANCE_TYPE_PSEUDONORMAL
// (Baerentzen & Aanaes 2005, angle-weighted pseudonormals):
//   FN  per_face_normals       unit (v1 - v0) x (v2 - v0), zero for a
//                              degenerate face (per_face_normals.cpp:19-31)
//   VN  per_vertex_normals     corner angles times FN summed at each vertex in
//       (ANGLE weighting)      face order, then normalised
//                              (per_vertex_normals.cpp:60-100; the angle
//                              2 atan(|u1 - u2| / |u1 + u2|) of unit edge
//                              vectors, internal_angles.cpp:20-38)
//   EN  per_edge_normals       FN summed over the faces of each undirected
//       (UNIFORM weighting)    edge in face order, then normalised
//                              (per_edge_normals.cpp:24-72)
// The sums run in the reference's loop order.  Rows of zero length are left
// as they are, as Eigen's normalize() leaves them.  EN is stored per face and
// corner (the edge opposite corner c of face f, the reference's EMAP(f + c*m)),
// so the sign test reads it without the edge map.
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <vector>

namespace gwn {

namespace {

inline double norm3(const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Eigen's normalize(): divide by the norm when the squared norm is positive
inline void normalize3(double* a) {
  const double z = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (z > 0.0) {
    const double r = std::sqrt(z);
    a[0] /= r;
    a[1] /= r;
    a[2] /= r;
  }
}

}  // namespace

void pseudonormals_host(const double* V, int64_t nv, const int32_t* F, int64_t nf, std::vector<double>& FN,
                        std::vector<double>& VN, std::vector<double>& ENf) {
  FN.assign(3 * (size_t)nf, 0.0);
  VN.assign(3 * (size_t)nv, 0.0);
  ENf.assign(9 * (size_t)nf, 0.0);
  for (int64_t f = 0; f < nf; ++f) {
    const double* p0 = V + 3 * (int64_t)F[3 * f];
    const double* p1 = V + 3 * (int64_t)F[3 * f + 1];
    const double* p2 = V + 3 * (int64_t)F[3 * f + 2];
    const double v1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double v2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double* n = &FN[3 * f];
    n[0] = v1[1] * v2[2] - v1[2] * v2[1];
    n[1] = v1[2] * v2[0] - v1[0] * v2[2];
    n[2] = v1[0] * v2[1] - v1[1] * v2[0];
    const double r = norm3(n);
    if (r == 0.0) {
      n[0] = n[1] = n[2] = 0.0;
    } else {
      n[0] /= r;
      n[1] /= r;
      n[2] /= r;
    }
  }
  // vertex normals: corner angle x face normal, faces in order, corners 0..2
  for (int64_t f = 0; f < nf; ++f) {
    for (int j = 0; j < 3; ++j) {
      const double* x = V + 3 * (int64_t)F[3 * f + (j + 2) % 3];
      const double* y = V + 3 * (int64_t)F[3 * f + j];
      const double* z = V + 3 * (int64_t)F[3 * f + (j + 1) % 3];
      double u1[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
      double u2[3] = {z[0] - y[0], z[1] - y[1], z[2] - y[2]};
      const double n1 = norm3(u1), n2 = norm3(u2);
      for (int k = 0; k < 3; ++k) {
        u1[k] /= n1;
        u2[k] /= n2;
      }
      const double dm[3] = {u1[0] - u2[0], u1[1] - u2[1], u1[2] - u2[2]};
      const double dp[3] = {u1[0] + u2[0], u1[1] + u2[1], u1[2] + u2[2]};
      const double w = 2.0 * std::atan(norm3(dm) / norm3(dp));
      double* vn = &VN[3 * (int64_t)F[3 * f + j]];
      for (int k = 0; k < 3; ++k) vn[k] += w * FN[3 * f + k];
    }
  }
  for (int64_t v = 0; v < nv; ++v) normalize3(&VN[3 * v]);
  // edge normals: the undirected edge opposite corner c of face f
  const size_t m = 3 * (size_t)nf;
  std::vector<uint64_t> key(m);
  std::vector<uint32_t> order(m);
  for (int64_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) {
      const uint32_t a = (uint32_t)F[3 * f + (c + 1) % 3], b = (uint32_t)F[3 * f + (c + 2) % 3];
      key[3 * f + c] = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
    }
  for (size_t k = 0; k < m; ++k) order[k] = (uint32_t)k;
  // stable: equal edges keep (f, c) order, which is the reference's
  // summation order (face-major, corners in order)
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return key[x] < key[y]; });
  for (size_t s = 0; s < m;) {
    size_t e = s;
    double sum[3] = {0.0, 0.0, 0.0};
    while (e < m && key[order[e]] == key[order[s]]) {
      const int64_t f = order[e] / 3;
      for (int k = 0; k < 3; ++k) sum[k] += FN[3 * f + k];
      ++e;
    }
    normalize3(sum);
    for (size_t t = s; t < e; ++t)
      for (int k = 0; k < 3; ++k) ENf[3 * (size_t)order[t] + k] = sum[k];
    s = e;
  }
}

}  // namespace gwn
