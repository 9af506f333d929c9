// soup_moments.h -- per-triangle and per-node Taylor moments of the soup path.
//
// Device restatement of UT_SolidAngle::init's precompute functors
// (FastWindingNumberForSoups.h:6644-7118).  T is the moment precision (float as
// in the reference, double for the all-FP64 form); boxes stay float (the
// positions' type S).  The operation order is the reference's, and the
// translation unit is compiled with -ffp-contract=off, so the FP32 BoxData
// equals the reference's bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gwn {

// LocalData (6612-6640): what a child hands to its parent.
template <typename T>
struct Local {
  float box[6];  // lo/hi per axis: x0 x1 y0 y1 z0 z1
  T P[3], AP[3], N[3];
  T area;
  T Nd[3];
  T Nxy, Nyx, Nyz, Nzy, Nzx, Nxz;
  T Nkd[3];
  T SP, Cxxy, Cxxz, Cyyz, Cyyx, Czzx, Czzy;
};

__device__ __forceinline__ float fmin_h(float a, float b) { return (a > b) ? b : a; }
__device__ __forceinline__ float fmax_h(float a, float b) { return (a < b) ? b : a; }
template <typename T>
__device__ __forceinline__ T tmax_h(T a, T b) {
  return (a < b) ? b : a;
}

// compute_integrals (6760-6844)
template <typename T>
__device__ __forceinline__ void integrals(const T* a, const T* b, const T* c, const T* P, T* iii, T* iij, T* iik,
                                          const int i) {
  const T oab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const T oac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const T ocb[3] = {b[0] - c[0], b[1] - c[1], b[2] - c[2]};
  const T t = oab[i] / oac[i];
  const int j = (i == 2) ? 0 : (i + 1);
  const int k = (j == 2) ? 0 : (j + 1);
  const T jdiff = t * oac[j] - oab[j];
  const T kdiff = t * oac[k] - oab[k];
  const T xa0 = (jdiff * oab[k] - kdiff * oab[j]), xa1 = kdiff * oab[i], xa2 = jdiff * oab[i];
  const T xc0 = (jdiff * ocb[k] - kdiff * ocb[j]), xc1 = kdiff * ocb[i], xc2 = jdiff * ocb[i];
  const T sa = sqrt(xa0 * xa0 + xa1 * xa1 + xa2 * xa2);
  const T sc = sqrt(xc0 * xc0 + xc1 * xc1 + xc2 * xc2);
  const T Pai = a[i] - P[i];
  const T Pci = c[i] - P[i];
  const T two3 = T(2.0 / 3.0), one3 = T(1.0 / 3.0), half = T(0.5), quarter = T(0.25);
  const T ia = sa * (half * Pai * Pai + two3 * Pai * oab[i] + quarter * oab[i] * oab[i]);
  const T ic = sc * (half * Pci * Pci + two3 * Pci * ocb[i] + quarter * ocb[i] * ocb[i]);
  *iii = ia + ic;
  int jk = j;
  T* out = iij;
  T diff = jdiff;
  for (int pass = 0; pass < 2; ++pass) {
    if (out) {
      const T obmid = b[jk] + half * diff;
      const T oabm = obmid - a[jk];
      const T ocbm = obmid - c[jk];
      const T Paj = a[jk] - P[jk];
      const T Pcj = c[jk] - P[jk];
      const T ja = sa * (half * Pai * Paj + one3 * Pai * oabm + one3 * Paj * oab[i] + quarter * oab[i] * oabm);
      const T jc = sc * (half * Pci * Pcj + one3 * Pci * ocbm + one3 * Pcj * ocb[i] + quarter * ocb[i] * ocbm);
      *out = ja + jc;
    }
    jk = k;
    out = iik;
    diff = kdiff;
  }
}

template <typename T>
__device__ __forceinline__ void sort3(const T* const v[3], int axis, int o[3]) {
  o[0] = 0;
  o[1] = 1;
  o[2] = 2;
  int t;
  if (v[0][axis] > v[1][axis]) {
    t = o[0]; o[0] = o[1]; o[1] = t;
  }
  if (v[o[0]][axis] > v[2][axis]) {
    t = o[0]; o[0] = o[2]; o[2] = t;
  }
  if (v[o[1]][axis] > v[o[2]][axis]) {
    t = o[1]; o[1] = o[2]; o[2] = t;
  }
}

// PrecomputeFunctors::item (6668-6890): one triangle's LocalData.
template <typename T>
__device__ void tri_local(const float* __restrict__ U, const int32_t* __restrict__ F,
                          const float* __restrict__ boxes, uint32_t itemi, int order, Local<T>& d) {
  const int32_t* tp = F + 3 * (int64_t)itemi;
  T a[3], b[3], c[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = (T)U[3 * (int64_t)tp[0] + k];
    b[k] = (T)U[3 * (int64_t)tp[1] + k];
    c[k] = (T)U[3 * (int64_t)tp[2] + k];
  }
  for (int k = 0; k < 6; ++k) d.box[k] = boxes[6 * (int64_t)itemi + k];
  const T ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const T ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const T half = T(0.5);
  const T N[3] = {half * (ab[1] * ac[2] - ab[2] * ac[1]), half * (ab[2] * ac[0] - ab[0] * ac[2]),
                  half * (ab[0] * ac[1] - ab[1] * ac[0])};
  const T area2 = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
  const T area = sqrt(area2);
  const T third = T(1) / T(3);
  T P[3];
  for (int k = 0; k < 3; ++k) {
    P[k] = ((a[k] + b[k]) + c[k]) * third;
    d.P[k] = P[k];
    d.AP[k] = P[k] * area;
    d.N[k] = N[k];
  }
  d.area = area;
  d.Nd[0] = d.Nd[1] = d.Nd[2] = T(0);
  d.Nxy = d.Nyx = d.Nyz = d.Nzy = d.Nzx = d.Nxz = T(0);
  d.Nkd[0] = d.Nkd[1] = d.Nkd[2] = T(0);
  d.SP = d.Cxxy = d.Cxxz = d.Cyyz = d.Cyyx = d.Czzx = d.Czzy = T(0);
  if (order < 2 || area == T(0)) return;
  const T ra = T(1) / area;
  const T n[3] = {N[0] * ra, N[1] * ra, N[2] * ra};
  const T* const vals[3] = {a, b, c};
  int ox[3], oy[3], oz[3];
  sort3<T>(vals, 0, ox);
  const T dx = vals[ox[2]][0] - vals[ox[0]][0];
  sort3<T>(vals, 1, oy);
  const T dy = vals[oy[2]][1] - vals[oy[0]][1];
  sort3<T>(vals, 2, oz);
  const T dz = vals[oz[2]][2] - vals[oz[0]][2];
  T ixx = 0, ixy = 0, iyy = 0, iyz = 0, izz = 0, izx = 0;
  if (dx > 0)
    integrals<T>(vals[ox[0]], vals[ox[1]], vals[ox[2]], P, &ixx, ((dx >= dy && dy > 0) ? &ixy : nullptr),
                 ((dx >= dz && dz > 0) ? &izx : nullptr), 0);
  if (dy > 0)
    integrals<T>(vals[oy[0]], vals[oy[1]], vals[oy[2]], P, &iyy, ((dy >= dz && dz > 0) ? &iyz : nullptr),
                 ((dx < dy && dx > 0) ? &ixy : nullptr), 1);
  if (dz > 0)
    integrals<T>(vals[oz[0]], vals[oz[1]], vals[oz[2]], P, &izz, ((dx < dz && dx > 0) ? &izx : nullptr),
                 ((dy < dz && dy > 0) ? &iyz : nullptr), 2);
  d.Nkd[0] = ixx * n[0];
  d.Nkd[1] = iyy * n[1];
  d.Nkd[2] = izz * n[2];
  d.SP = T(2) * (n[0] * iyz + n[1] * izx + n[2] * ixy);
  d.Cxxy = T(2) * (n[0] * ixy) + n[1] * ixx;
  d.Cxxz = T(2) * (n[0] * izx) + n[2] * ixx;
  d.Cyyz = T(2) * (n[1] * iyz) + n[2] * iyy;
  d.Cyyx = T(2) * (n[1] * ixy) + n[0] * iyy;
  d.Czzx = T(2) * (n[2] * izx) + n[0] * izz;
  d.Czzy = T(2) * (n[2] * iyz) + n[1] * izz;
}

// BoxData for one node in the reference member order (6477-6521), 92 values:
// [0..3] maxP2, [4..15] P (x4,y4,z4), [16..27] N, [28..39] NijDiag,
// [40..43] Nxy+Nyx, [44..47] Nyz+Nzy, [48..51] Nzx+Nxz, [52..63] NijkDiag,
// [64..67] SumPermuteNxyz, [68..91] 2Nxxy+Nyxx, 2Nxxz+Nzxx, 2Nyyz+Nzyy,
// 2Nyyx+Nxyy, 2Nzzx+Nxzz, 2Nzzy+Nyzz.
enum : int {
  BD_MAXP2 = 0, BD_P = 4, BD_N = 16, BD_ND = 28, BD_NXY = 40, BD_NYZ = 44, BD_NZX = 48, BD_NKD = 52,
  BD_SP = 64, BD_CXXY = 68, BD_CXXZ = 72, BD_CYYZ = 76, BD_CYYX = 80, BD_CZZX = 84, BD_CZZY = 88
};

// PrecomputeFunctors::post (6892-7118): combine children into the node.
template <typename T>
__device__ void node_post(const Local<T>* ch, int nch, int order, T* bd, Local<T>& out) {
  T N[3] = {ch[0].N[0], ch[0].N[1], ch[0].N[2]};
  for (int k = 0; k < 3; ++k) {
    bd[BD_N + 4 * k] = N[k];
    bd[BD_P + 4 * k] = ch[0].P[k];
  }
  T AP[3] = {ch[0].AP[0], ch[0].AP[1], ch[0].AP[2]};
  T area = ch[0].area;
  for (int i = 1; i < nch; ++i) {
    for (int k = 0; k < 3; ++k) {
      N[k] += ch[i].N[k];
      bd[BD_N + 4 * k + i] = ch[i].N[k];
      AP[k] += ch[i].AP[k];
      bd[BD_P + 4 * k + i] = ch[i].P[k];
    }
    area += ch[i].area;
  }
  for (int i = nch; i < 4; ++i)
    for (int k = 0; k < 3; ++k) bd[BD_N + 4 * k + i] = bd[BD_P + 4 * k + i] = T(0);
  for (int k = 0; k < 3; ++k) {
    out.N[k] = N[k];
    out.AP[k] = AP[k];
  }
  out.area = area;
  float box[6];
  for (int k = 0; k < 6; ++k) box[k] = ch[0].box[k];
  for (int i = 1; i < nch; ++i)
    for (int a = 0; a < 3; ++a) {
      box[2 * a] = (box[2 * a] < ch[i].box[2 * a]) ? box[2 * a] : ch[i].box[2 * a];
      box[2 * a + 1] = (box[2 * a + 1] > ch[i].box[2 * a + 1]) ? box[2 * a + 1] : ch[i].box[2 * a + 1];
    }
  if (area > T(0)) {
    const T r = T(1) / area;
    for (int k = 0; k < 3; ++k) out.P[k] = AP[k] * r;
  } else {
    for (int k = 0; k < 3; ++k) out.P[k] = T(0.5) * ((T)box[2 * k] + (T)box[2 * k + 1]);
  }
  for (int k = 0; k < 6; ++k) out.box[k] = box[k];
  for (int i = 0; i < nch; ++i) {
    T dd[3];
    for (int k = 0; k < 3; ++k)
      dd[k] = tmax_h<T>(ch[i].P[k] - (T)ch[i].box[2 * k], (T)ch[i].box[2 * k + 1] - ch[i].P[k]);
    bd[BD_MAXP2 + i] = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
  }
  for (int i = nch; i < 4; ++i) bd[BD_MAXP2 + i] = (T)INFINITY;

  // order >= 1 terms (the reference copies them for order 1 as well)
  for (int k = 0; k < 3; ++k) {
    out.Nd[k] = ch[0].Nd[k];
    out.Nkd[k] = ch[0].Nkd[k];
  }
  out.Nxy = out.Nyx = out.Nyz = out.Nzy = out.Nzx = out.Nxz = T(0);
  out.SP = ch[0].SP;
  out.Cxxy = ch[0].Cxxy;
  out.Cxxz = ch[0].Cxxz;
  out.Cyyz = ch[0].Cyyz;
  out.Cyyx = ch[0].Cyyx;
  out.Czzx = ch[0].Czzx;
  out.Czzy = ch[0].Czzy;
  for (int i = 1; i < nch; ++i) {
    for (int k = 0; k < 3; ++k) {
      out.Nd[k] += ch[i].Nd[k];
      out.Nkd[k] += ch[i].Nkd[k];
    }
    out.SP += ch[i].SP;
    out.Cxxy += ch[i].Cxxy;
    out.Cxxz += ch[i].Cxxz;
    out.Cyyz += ch[i].Cyyz;
    out.Cyyx += ch[i].Cyyx;
    out.Czzx += ch[i].Czzx;
    out.Czzy += ch[i].Czzy;
  }
  for (int i = 0; i < 4; ++i) {
    const bool on = i < nch;
    for (int k = 0; k < 3; ++k) {
      bd[BD_ND + 4 * k + i] = on ? ch[i].Nd[k] : T(0);
      bd[BD_NKD + 4 * k + i] = on ? ch[i].Nkd[k] : T(0);
    }
    bd[BD_NXY + i] = on ? ch[i].Nxy + ch[i].Nyx : T(0);
    bd[BD_NYZ + i] = on ? ch[i].Nyz + ch[i].Nzy : T(0);
    bd[BD_NZX + i] = on ? ch[i].Nzx + ch[i].Nxz : T(0);
    bd[BD_SP + i] = on ? ch[i].SP : T(0);
    bd[BD_CXXY + i] = on ? ch[i].Cxxy : T(0);
    bd[BD_CXXZ + i] = on ? ch[i].Cxxz : T(0);
    bd[BD_CYYZ + i] = on ? ch[i].Cyyz : T(0);
    bd[BD_CYYX + i] = on ? ch[i].Cyyx : T(0);
    bd[BD_CZZX + i] = on ? ch[i].Czzx : T(0);
    bd[BD_CZZY + i] = on ? ch[i].Czzy : T(0);
  }
  for (int i = 0; i < nch; ++i) {
    const Local<T>& c = ch[i];
    T dp[3];
    for (int k = 0; k < 3; ++k) dp[k] = c.P[k] - out.P[k];
    const T* n = c.N;
    for (int k = 0; k < 3; ++k) out.Nd[k] += n[k] * dp[k];
    const T Nxy = c.Nxy + n[0] * dp[1];
    const T Nyx = c.Nyx + n[1] * dp[0];
    const T Nyz = c.Nyz + n[1] * dp[2];
    const T Nzy = c.Nzy + n[2] * dp[1];
    const T Nzx = c.Nzx + n[2] * dp[0];
    const T Nxz = c.Nxz + n[0] * dp[2];
    out.Nxy += Nxy;
    out.Nyx += Nyx;
    out.Nyz += Nyz;
    out.Nzy += Nzy;
    out.Nzx += Nzx;
    out.Nxz += Nxz;
    if (order >= 2) {
      const T two = T(2);
      for (int k = 0; k < 3; ++k) out.Nkd[k] += (two * dp[k]) * c.Nd[k] + (dp[k] * dp[k]) * n[k];
      out.SP += (dp[0] * (Nyz + Nzy) + dp[1] * (Nzx + Nxz) + dp[2] * (Nxy + Nyx));
      out.Cxxy += two * (dp[1] * c.Nd[0] + dp[0] * c.Nxy + n[0] * dp[0] * dp[1]) + two * c.Nyx * dp[0] +
                  n[1] * dp[0] * dp[0];
      out.Cxxz += two * (dp[2] * c.Nd[0] + dp[0] * c.Nxz + n[0] * dp[0] * dp[2]) + two * c.Nzx * dp[0] +
                  n[2] * dp[0] * dp[0];
      out.Cyyz += two * (dp[2] * c.Nd[1] + dp[1] * c.Nyz + n[1] * dp[1] * dp[2]) + two * c.Nzy * dp[1] +
                  n[2] * dp[1] * dp[1];
      out.Cyyx += two * (dp[0] * c.Nd[1] + dp[1] * c.Nyx + n[1] * dp[1] * dp[0]) + two * c.Nxy * dp[1] +
                  n[0] * dp[1] * dp[1];
      out.Czzx += two * (dp[0] * c.Nd[2] + dp[2] * c.Nzx + n[2] * dp[2] * dp[0]) + two * c.Nxz * dp[2] +
                  n[0] * dp[2] * dp[2];
      out.Czzy += two * (dp[1] * c.Nd[2] + dp[2] * c.Nzy + n[2] * dp[2] * dp[1]) + two * c.Nyz * dp[2] +
                  n[1] * dp[2] * dp[2];
    }
  }
}

}  // namespace gwn
