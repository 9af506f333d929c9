/*
 * oracle_distance.c -- CPU restatement of igl::signed_distance for the
 * pseudonormal, unsigned, winding-number and fast-winding-number sign types.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   oracle_point_tri_sqrdist   point_simplex_squared_distance.cpp:43-114 (the
 *                              closest-point-on-triangle method of Ericson,
 *                              Real-Time Collision Detection, sec. 5.1.5)
 *   oracle_signed_distance     signed_distance.cpp:28-140 with the AABB
 *                              query replaced by a brute-force minimum over
 *                              all faces (same strict "<" replacement rule as
 *                              AABB.cpp set_min, faces in index order)
 *   oracle_pseudonormals       per_face_normals.cpp:19-31,
 *                              per_vertex_normals.cpp:60-100 (ANGLE weights,
 *                              internal_angles.cpp:20-38),
 *                              per_edge_normals.cpp:24-72 (UNIFORM weights)
 *   oracle_pseudonormal_at     pseudonormal_test.cpp:24-141 (doublearea.cpp:
 *                              144-197 Kahan/Heron, barycentric_coordinates.cpp:
 *                              88-100, project_to_line.cpp / _segment.cpp)
 *   oracle_winding_number_2d   winding_number.cpp:62-89 for 2D curves
 *                              (signed_angle.cpp:8-41 summed over edges)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

double oracle_point_tri_sqrdist(const double* p, const double* a, const double* b, const double* c, double* out) {
  double ab[3], ac[3], ap[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = b[k] - a[k];
    ac[k] = c[k] - a[k];
    ap[k] = p[k] - a[k];
  }
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  double q[3];
  /* vertex region of a (point_simplex_squared_distance.cpp:49-58) */
  if (d1 <= 0.0 && d2 <= 0.0) {
    memcpy(q, a, sizeof q);
    goto done;
  }
  double bp[3];
  for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  /* vertex region of b (60-67) */
  if (d3 >= 0.0 && d4 <= d3) {
    memcpy(q, b, sizeof q);
    goto done;
  }
  /* edge ab (68-78), skipped for a degenerate edge a == b */
  const double vc = d1 * d4 - d3 * d2;
  if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) {
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
      const double v = d1 / (d1 - d3);
      for (int k = 0; k < 3; ++k) q[k] = a[k] + v * ab[k];
      goto done;
    }
  }
  double cp[3];
  for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  /* vertex region of c (79-87) */
  if (d6 >= 0.0 && d5 <= d6) {
    memcpy(q, c, sizeof q);
    goto done;
  }
  /* edge ac (88-95) */
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    for (int k = 0; k < 3; ++k) q[k] = a[k] + w * ac[k];
    goto done;
  }
  /* edge bc (96-103) */
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; ++k) q[k] = b[k] + w * (c[k] - b[k]);
    goto done;
  }
  /* face interior (104-110) */
  {
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    for (int k = 0; k < 3; ++k) q[k] = a[k] + ab[k] * v + ac[k] * w;
  }
done:
  if (out) memcpy(out, q, sizeof q);
  {
    const double e0 = p[0] - q[0], e1 = p[1] - q[1], e2 = p[2] - q[2];
    return e0 * e0 + e1 * e1 + e2 * e2;
  }
}

typedef struct {
  const double* V;
  const int32_t* F;
  int64_t nf;
  const double* P;
  double* d2;
  int32_t* I;
  double* C;
  double up2;
} sd_ctx;

static void sd_range(int64_t b, int64_t e, void* vctx) {
  const sd_ctx* x = (const sd_ctx*)vctx;
  for (int64_t p = b; p < e; ++p) {
    const double* q = x->P + 3 * p;
    double best = x->up2, bc[3] = {0, 0, 0};
    int32_t bi = -1;
    for (int64_t f = 0; f < x->nf; ++f) {
      const int32_t* t = x->F + 3 * f;
      double c[3];
      const double d2 = oracle_point_tri_sqrdist(q, x->V + 3 * (int64_t)t[0], x->V + 3 * (int64_t)t[1],
                                                 x->V + 3 * (int64_t)t[2], c);
      if (d2 < best) {
        best = d2;
        bi = (int32_t)f;
        memcpy(bc, c, sizeof bc);
      }
    }
    x->d2[p] = best;
    x->I[p] = bi;
    memcpy(x->C + 3 * p, bc, sizeof bc);
  }
}

static double std_max(double a, double b) { return (a < b) ? b : a; }

static double nrm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

/* Eigen's normalize(): unchanged when the squared norm is not positive */
static void unit3(double* a) {
  const double z = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (z > 0.0) {
    const double r = sqrt(z);
    a[0] /= r;
    a[1] /= r;
    a[2] /= r;
  }
}

typedef struct {
  uint64_t key;
  int64_t slot; /* 3 * f + c */
} edge_slot;

static int edge_slot_cmp(const void* pa, const void* pb) {
  const edge_slot* a = (const edge_slot*)pa;
  const edge_slot* b = (const edge_slot*)pb;
  if (a->key != b->key) return a->key < b->key ? -1 : 1;
  return (a->slot > b->slot) - (a->slot < b->slot);
}

int oracle_pseudonormals(const double* V, int64_t nv, const int32_t* F, int64_t nf, double* FN, double* VN,
                         double* EN) {
  for (int64_t f = 0; f < nf; ++f) { /* per_face_normals.cpp:19-31 */
    const double* a = V + 3 * (int64_t)F[3 * f];
    const double* b = V + 3 * (int64_t)F[3 * f + 1];
    const double* c = V + 3 * (int64_t)F[3 * f + 2];
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double* n = FN + 3 * f;
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double r = nrm3(n[0], n[1], n[2]);
    if (r == 0.0)
      n[0] = n[1] = n[2] = 0.0;
    else
      for (int k = 0; k < 3; ++k) n[k] /= r;
  }
  for (int64_t i = 0; i < 3 * nv; ++i) VN[i] = 0.0;
  for (int64_t f = 0; f < nf; ++f) /* per_vertex_normals.cpp:87-96, corner angles internal_angles.cpp:20-38 */
    for (int j = 0; j < 3; ++j) {
      const double* x = V + 3 * (int64_t)F[3 * f + (j + 2) % 3];
      const double* y = V + 3 * (int64_t)F[3 * f + j];
      const double* z = V + 3 * (int64_t)F[3 * f + (j + 1) % 3];
      double v1[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]}, v2[3] = {z[0] - y[0], z[1] - y[1], z[2] - y[2]};
      const double l1 = nrm3(v1[0], v1[1], v1[2]), l2 = nrm3(v2[0], v2[1], v2[2]);
      for (int k = 0; k < 3; ++k) {
        v1[k] /= l1;
        v2[k] /= l2;
      }
      const double ang = 2.0 * atan(nrm3(v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]) /
                                    nrm3(v1[0] + v2[0], v1[1] + v2[1], v1[2] + v2[2]));
      double* vn = VN + 3 * (int64_t)F[3 * f + j];
      for (int k = 0; k < 3; ++k) vn[k] += ang * FN[3 * f + k];
    }
  for (int64_t v = 0; v < nv; ++v) unit3(VN + 3 * v);
  /* per_edge_normals.cpp:24-72: FN summed per undirected edge, (f, c) order */
  edge_slot* es = (edge_slot*)malloc(sizeof(edge_slot) * (size_t)(nf > 0 ? 3 * nf : 1));
  if (!es) return -2;
  for (int64_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) {
      const uint32_t a = (uint32_t)F[3 * f + (c + 1) % 3], b = (uint32_t)F[3 * f + (c + 2) % 3];
      es[3 * f + c].key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
      es[3 * f + c].slot = 3 * f + c;
    }
  qsort(es, (size_t)(3 * nf), sizeof(edge_slot), edge_slot_cmp);
  for (int64_t s0 = 0; s0 < 3 * nf;) {
    int64_t s1 = s0;
    double sum[3] = {0.0, 0.0, 0.0};
    while (s1 < 3 * nf && es[s1].key == es[s0].key) {
      const int64_t f = es[s1].slot / 3;
      for (int k = 0; k < 3; ++k) sum[k] += FN[3 * f + k];
      ++s1;
    }
    unit3(sum);
    for (int64_t t = s0; t < s1; ++t)
      for (int k = 0; k < 3; ++k) EN[3 * es[t].slot + k] = sum[k];
    s0 = s1;
  }
  free(es);
  return 0;
}

void oracle_pseudonormal_at(const double* V, const int32_t* F, const double* FN, const double* VN, const double* EN,
                            int64_t f, const double* c, double* n) {
  const double eps = 1e-12;
  const int32_t* t = F + 3 * f;
  const double* A = V + 3 * (int64_t)t[0];
  const double* B = V + 3 * (int64_t)t[1];
  const double* Cc = V + 3 * (int64_t)t[2];
  double l[3] = {nrm3(B[0] - Cc[0], B[1] - Cc[1], B[2] - Cc[2]), nrm3(Cc[0] - A[0], Cc[1] - A[1], Cc[2] - A[2]),
                 nrm3(A[0] - B[0], A[1] - B[1], A[2] - B[2])};
  /* igl::sort(ul, 2, false): descending */
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (l[j] > l[i]) {
        const double x = l[i];
        l[i] = l[j];
        l[j] = x;
      }
  const double arg = (l[0] + (l[1] + l[2])) * (l[2] - (l[0] - l[1])) * (l[2] + (l[0] - l[1])) * (l[0] + (l[1] - l[2]));
  const double area = 2.0 * 0.25 * sqrt(arg);
  const double* src = FN + 3 * f;
  if (area > 1e-4) {
    const double v0[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const double v1[3] = {Cc[0] - A[0], Cc[1] - A[1], Cc[2] - A[2]};
    const double v2[3] = {c[0] - A[0], c[1] - A[1], c[2] - A[2]};
    const double d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    const double den = d00 * d11 - d01 * d01;
    double b[3];
    b[1] = (d11 * d20 - d01 * d21) / den;
    b[2] = (d00 * d21 - d01 * d20) / den;
    b[0] = 1.0 - (b[1] + b[2]);
    const int type = (b[0] <= eps) + (b[1] <= eps) + (b[2] <= eps);
    if (type == 2) {
      for (int x = 0; x < 3; ++x)
        if (b[x] > eps) {
          src = VN + 3 * (int64_t)t[x];
          break;
        }
    } else if (type == 1) {
      for (int x = 0; x < 3; ++x)
        if (b[x] <= eps) {
          src = EN + 9 * f + 3 * x;
          break;
        }
    }
  } else {
    int found = 0;
    for (int v = 0; v < 3 && !found; ++v) {
      const double* Pv = V + 3 * (int64_t)t[v];
      if (nrm3(c[0] - Pv[0], c[1] - Pv[1], c[2] - Pv[2]) < eps) {
        found = 1;
        src = VN + 3 * (int64_t)t[v];
      }
    }
    for (int e = 0; e < 3 && !found; ++e) {
      const double* Sv = V + 3 * (int64_t)t[(e + 1) % 3];
      const double* Dv = V + 3 * (int64_t)t[(e + 2) % 3];
      const double dms[3] = {Dv[0] - Sv[0], Dv[1] - Sv[1], Dv[2] - Sv[2]};
      const double smp[3] = {Sv[0] - c[0], Sv[1] - c[1], Sv[2] - c[2]};
      const double tt = -dot3(dms, smp) / dot3(dms, dms);
      const double pr[3] = {(1 - tt) * Sv[0] + tt * Dv[0], (1 - tt) * Sv[1] + tt * Dv[1], (1 - tt) * Sv[2] + tt * Dv[2]};
      double sq = (c[0] - pr[0]) * (c[0] - pr[0]) + (c[1] - pr[1]) * (c[1] - pr[1]) + (c[2] - pr[2]) * (c[2] - pr[2]);
      if (tt < 0)
        sq = (c[0] - Sv[0]) * (c[0] - Sv[0]) + (c[1] - Sv[1]) * (c[1] - Sv[1]) + (c[2] - Sv[2]) * (c[2] - Sv[2]);
      else if (tt > 1)
        sq = (c[0] - Dv[0]) * (c[0] - Dv[0]) + (c[1] - Dv[1]) * (c[1] - Dv[1]) + (c[2] - Dv[2]) * (c[2] - Dv[2]);
      if (sqrt(sq) < eps) {
        src = EN + 9 * f + 3 * e;
        found = 1;
      }
    }
  }
  n[0] = src[0];
  n[1] = src[1];
  n[2] = src[2];
}

int oracle_signed_distance(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* P, int64_t np,
                           int sign_type, double lower, double upper, double* S, int32_t* I, double* C,
                           int nthreads) {
  return oracle_signed_distance_n(V, nv, F, nf, P, np, sign_type, lower, upper, S, I, C, NULL, nthreads);
}

int oracle_signed_distance_n(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* P, int64_t np,
                             int sign_type, double lower, double upper, double* S, int32_t* I, double* C, double* N,
                             int nthreads) {
  if (sign_type < 0 || sign_type > 4) return -1;
  if (nf == 0) { /* signed_distance.cpp:397-403 */
    for (int64_t p = 0; p < np; ++p) {
      S[p] = NAN;
      I[p] = -1;
      C[3 * p] = C[3 * p + 1] = C[3 * p + 2] = NAN;
    }
    return 0;
  }
  /* signed_distance.cpp:79-84 */
  const double max_abs = std_max(fabs(lower), fabs(upper));
  const double up2 = pow(max_abs, 2.0);
  const double low2 = pow(std_max(max_abs - (upper - lower), 0.0), 2.0);
  double* d2 = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  double* W = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  if (!d2 || !W) {
    free(d2);
    free(W);
    return -2;
  }
  sd_ctx x = {V, F, nf, P, d2, I, C, up2};
  oracle_parallel_for(np, sd_range, &x, 1, nthreads);
  if (sign_type == 4) {
    /* fast_winding_number(V.cast<float>(), F, 2, fwn_bvh) then the
     * single-point query at accuracy_scale 2 (fast_winding_number.cpp:439-455) */
    float* U = (float*)malloc(sizeof(float) * 3 * (size_t)nv);
    if (!U) {
      free(d2);
      free(W);
      return -2;
    }
    for (int64_t i = 0; i < 3 * nv; ++i) U[i] = (float)V[i];
    oracle_soup* s = oracle_soup_build(U, (int32_t)nv, F, (int32_t)nf, 2);
    for (int64_t p = 0; p < np; ++p) {
      const float q[3] = {(float)P[3 * p], (float)P[3 * p + 1], (float)P[3 * p + 2]};
      const float w = (float)(oracle_soup_solid_angle(s, q, 2.0f) / (4.0 * 3.14159265358979323846));
      W[p] = (double)w;
    }
    oracle_soup_free(s);
    free(U);
  } else if (sign_type == 1 || sign_type == 2) {
    oracle_winding_number_bf(V, nv, F, nf, P, np, W, nthreads);
  }
  double *FN = NULL, *VN = NULL, *EN = NULL;
  if (sign_type == 0) { /* signed_distance.cpp:68-76 */
    FN = (double*)malloc(sizeof(double) * 3 * (size_t)nf);
    VN = (double*)malloc(sizeof(double) * 3 * (size_t)(nv > 0 ? nv : 1));
    EN = (double*)malloc(sizeof(double) * 9 * (size_t)nf);
    if (!FN || !VN || !EN || oracle_pseudonormals(V, nv, F, nf, FN, VN, EN) != 0) {
      free(FN);
      free(VN);
      free(EN);
      free(d2);
      free(W);
      return -2;
    }
  }
  for (int64_t p = 0; p < np; ++p) {
    if (d2[p] >= up2 || d2[p] < low2) { /* signed_distance.cpp:96-103 */
      S[p] = NAN;
      I[p] = (int32_t)(nf + 1);
      C[3 * p] = C[3 * p + 1] = C[3 * p + 2] = 0.0;
      if (N && sign_type == 0) N[3 * p] = N[3 * p + 1] = N[3 * p + 2] = 0.0;
      continue;
    }
    double s = 1.0;
    if (sign_type == 4)
      s = 1.0 - 2.0 * fabs(W[p]);
    else if (sign_type == 1 || sign_type == 2)
      s = 1.0 - 2.0 * W[p];
    else if (sign_type == 0) { /* signed_distance.cpp:125-128 */
      double n[3];
      const double* c = C + 3 * p;
      oracle_pseudonormal_at(V, F, FN, VN, EN, I[p], c, n);
      const double qc[3] = {P[3 * p] - c[0], P[3 * p + 1] - c[1], P[3 * p + 2] - c[2]};
      s = dot3(qc, n) >= 0 ? 1.0 : -1.0;
      if (N) memcpy(N + 3 * p, n, sizeof n);
    }
    S[p] = s * sqrt(d2[p]);
  }
  free(FN);
  free(VN);
  free(EN);
  free(d2);
  free(W);
  return 0;
}

/* ---- 2D curves: winding_number.cpp:62-89 (F.cols() == 2) summing
 * signed_angle.cpp:8-41 over edges in order ---- */
double oracle_signed_angle_2d(const double* a, const double* b, const double* p) {
  double oa[2], ob[2], la = 0, lb = 0;
  for (int i = 0; i < 2; ++i) {
    oa[i] = p[i] - a[i];
    ob[i] = p[i] - b[i];
    la += oa[i] * oa[i];
    lb += ob[i] * ob[i];
  }
  la = sqrt(la);
  lb = sqrt(lb);
  for (int i = 0; i < 2; ++i) {
    if (la != 0) oa[i] /= la;
    if (lb != 0) ob[i] /= lb;
  }
  return -atan2(ob[0] * oa[1] - ob[1] * oa[0], ob[0] * oa[0] + ob[1] * oa[1]) / (2. * 3.1415926535897932384626433832795);
}

void oracle_winding_number_2d(const double* V, const int32_t* E, int64_t ne, const double* O, int64_t no,
                              double* W) {
  for (int64_t o = 0; o < no; ++o) {
    double w = 0;
    for (int64_t e = 0; e < ne; ++e) w += oracle_signed_angle_2d(V + 2 * (int64_t)E[2 * e], V + 2 * (int64_t)E[2 * e + 1], O + 2 * o);
    W[o] = w;
  }
}
