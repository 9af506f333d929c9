/*
 * oracle_distance.c -- CPU restatement of igl::signed_distance for the
 * unsigned, winding-number and fast-winding-number sign types.
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

int oracle_signed_distance(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* P, int64_t np,
                           int sign_type, double lower, double upper, double* S, int32_t* I, double* C,
                           int nthreads) {
  if (sign_type != 1 && sign_type != 2 && sign_type != 3 && sign_type != 4) return -1;
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
  for (int64_t p = 0; p < np; ++p) {
    if (d2[p] >= up2 || d2[p] < low2) { /* signed_distance.cpp:96-103 */
      S[p] = NAN;
      I[p] = (int32_t)(nf + 1);
      C[3 * p] = C[3 * p + 1] = C[3 * p + 2] = 0.0;
      continue;
    }
    double s = 1.0;
    if (sign_type == 4)
      s = 1.0 - 2.0 * fabs(W[p]);
    else if (sign_type == 1 || sign_type == 2)
      s = 1.0 - 2.0 * W[p];
    S[p] = s * sqrt(d2[p]);
  }
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
