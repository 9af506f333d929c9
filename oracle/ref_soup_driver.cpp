// ref_soup_driver.cpp -- TEST INFRASTRUCTURE (the checker, never the product).
//
// A C-callable wrapper around the reference's own UT_SolidAngle, compiled from
// include/igl/FastWindingNumberForSoups.h where it lies in the reference tree
// (see oracle/Makefile).  The output library goes to oracle/_ref/ only.  It
// exposes the reference tree (BVH<4> nodes, 4217-4232), the per-node BoxData
// (6477-6521) and computeSolidAngle (7148-7284), and the same for the 2D
// UT_SubtendedAngle (BoxData 7287-7320, computeAngle 7700-7828), so that the C
// restatements in oracle_soup.c / oracle_curve.c and the HIP path can be
// compared with the reference itself.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <iostream>
#include <memory>
#include <thread>
#include <vector>
// BoxData and the tree are private members of UT_SolidAngle; widen access for
// this test wrapper only (the reference sources are used unmodified).
#define private public
#include <igl/FastWindingNumberForSoups.h>
#undef private

using namespace igl::FastWindingNumber::HDK_Sample;

namespace {
struct RefSoup {
  std::vector<UT_Vector3T<float>> U;
  std::vector<int> F;
  UT_SolidAngle<float, float> sa;
};
}  // namespace

extern "C" {

void* ref_soup_build(const float* V, int nv, const int* F, int nf, int order) {
  RefSoup* r = new RefSoup();
  r->U.resize(nv > 0 ? nv : 1);
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < 3; ++j) r->U[i][j] = V[3 * i + j];
  r->F.assign(F, F + 3 * (size_t)nf);
  if (r->F.empty()) r->F.push_back(0);
  r->sa.init(nf, r->F.data(), nv, r->U.data(), order);
  return r;
}

void ref_soup_free(void* h) { delete static_cast<RefSoup*>(h); }

int ref_soup_num_nodes(void* h) { return static_cast<RefSoup*>(h)->sa.myTree.getNumNodes(); }

void ref_soup_nodes(void* h, uint32_t* out) {
  RefSoup* r = static_cast<RefSoup*>(h);
  const int n = r->sa.myTree.getNumNodes();
  if (n) std::memcpy(out, r->sa.myTree.getNodes(), sizeof(uint32_t) * 4 * (size_t)n);
}

void ref_soup_boxdata(void* h, float* out) {
  RefSoup* r = static_cast<RefSoup*>(h);
  const int n = r->sa.myTree.getNumNodes();
  static_assert(sizeof(*r->sa.myData.get()) == 92 * sizeof(float), "BoxData layout");
  if (n) std::memcpy(out, r->sa.myData.get(), sizeof(float) * 92 * (size_t)n);
}

// Omega for each query (float), serial; the caller divides by 4 pi.
void ref_soup_solid_angle(void* h, const float* Q, int64_t nq, float accuracy_scale, float* omega) {
  RefSoup* r = static_cast<RefSoup*>(h);
  for (int64_t p = 0; p < nq; ++p) {
    UT_Vector3T<float> q;
    q[0] = Q[3 * p];
    q[1] = Q[3 * p + 1];
    q[2] = Q[3 * p + 2];
    omega[p] = r->sa.computeSolidAngle(q, accuracy_scale);
  }
}

// fast_winding_number(bvh, accuracy_scale, Q, W) (fast_winding_number.cpp:418-437)
// with the reference's own igl::parallel_for; Q double, W = Omega/(4 pi).
void ref_soup_query(void* h, const double* Q, int64_t nq, float accuracy_scale, double* W) {
  RefSoup* r = static_cast<RefSoup*>(h);
  igl::parallel_for(
      (int)nq,
      [&](int p) {
        UT_Vector3T<float> q;
        q[0] = (float)Q[3 * (int64_t)p];
        q[1] = (float)Q[3 * (int64_t)p + 1];
        q[2] = (float)Q[3 * (int64_t)p + 2];
        W[p] = r->sa.computeSolidAngle(q, accuracy_scale) / (4.0 * 3.1415926535897932384626433832795);
      },
      1000);
}

// UTsignedSolidAngleTri (6070-6110) instantiated for double.
double ref_solid_angle_tri_f64(const double* a, const double* b, const double* c, const double* q) {
  UT_Vector3T<double> A, B, C, Qp;
  for (int i = 0; i < 3; ++i) {
    A[i] = a[i];
    B[i] = b[i];
    C[i] = c[i];
    Qp[i] = q[i];
  }
  return UTsignedSolidAngleTri(A, B, C, Qp);
}

int ref_num_threads(void) { return (int)igl::default_num_threads(); }

// ---- UT_SubtendedAngle<float,float> (6315-6370, init 7342-7604, computeAngle 7700-7828) ----
struct RefCurve {
  std::vector<UT_Vector2T<float>> P;
  std::vector<int> E;
  UT_SubtendedAngle<float, float> sa;
};

void* ref_curve_build(const float* P, int np, const int* E, int ne, int order) {
  RefCurve* r = new RefCurve();
  r->P.resize(np > 0 ? np : 1);
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < 2; ++j) r->P[i][j] = P[2 * i + j];
  r->E.assign(E, E + 2 * (size_t)ne);
  if (r->E.empty()) r->E.push_back(0);
  r->sa.init(ne, r->E.data(), np, r->P.data(), order);
  return r;
}

void ref_curve_free(void* h) { delete static_cast<RefCurve*>(h); }

int ref_curve_num_nodes(void* h) { return static_cast<RefCurve*>(h)->sa.myTree.getNumNodes(); }

void ref_curve_nodes(void* h, uint32_t* out) {
  RefCurve* r = static_cast<RefCurve*>(h);
  const int n = r->sa.myTree.getNumNodes();
  if (n) std::memcpy(out, r->sa.myTree.getNodes(), sizeof(uint32_t) * 4 * (size_t)n);
}

// BoxData (7287-7320): 12 x v4uf = 48 floats per node
void ref_curve_boxdata(void* h, float* out) {
  RefCurve* r = static_cast<RefCurve*>(h);
  const int n = r->sa.myTree.getNumNodes();
  static_assert(sizeof(*r->sa.myData.get()) == 48 * sizeof(float), "BoxData layout");
  if (n) std::memcpy(out, r->sa.myData.get(), sizeof(float) * 48 * (size_t)n);
}

// computeAngle for each query (float), serial
void ref_curve_angle(void* h, const float* Q, int64_t nq, float accuracy_scale, float* angle) {
  RefCurve* r = static_cast<RefCurve*>(h);
  for (int64_t p = 0; p < nq; ++p) {
    UT_Vector2T<float> q;
    q[0] = Q[2 * p];
    q[1] = Q[2 * p + 1];
    angle[p] = r->sa.computeAngle(q, accuracy_scale);
  }
}

// computeAngle for each query over the host cores with the reference's own
// igl::parallel_for (min_parallel 1000, as its batch soup query uses)
void ref_curve_angle_par(void* h, const float* Q, int64_t nq, float accuracy_scale, float* angle) {
  RefCurve* r = static_cast<RefCurve*>(h);
  igl::parallel_for(
      (int)nq,
      [&](int p) {
        UT_Vector2T<float> q;
        q[0] = Q[2 * (int64_t)p];
        q[1] = Q[2 * (int64_t)p + 1];
        angle[p] = r->sa.computeAngle(q, accuracy_scale);
      },
      1000);
}
}
