/*
 * oracle_antipodal.c -- CPU restatement of the antipodal-method generalized
 * winding number (WindingNumberAntipodalScene.h:48-256,
 * winding_number_antipodal.cpp).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   boundary    build_boundary_segments (WindingNumberAntipodalScene.h:
 *               205-253): +1 per triangle traversing an undirected edge
 *               min->max, -1 for max->min; non-zero nets become segments
 *               oriented so the weight is positive.  Edges are taken in
 *               (min, max) order (the reference's unordered_map order is
 *               unspecified).
 *   frac        half_solid_angle_unorm (180-191) summed over the segments,
 *               divided by 2 pi (90-97)
 *   crossings   the signed count the reference asks of its intersector
 *               (embree/EmbreeIntersector.cpp:390-443: +1 where the ray leaves
 *               through a triangle's front), by brute force over every
 *               triangle with the watertight ray-triangle test of Woop,
 *               Benthin and Wald (JCGT 2013) in FP64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
  uint64_t key;
  int sgn;
} ap_edge;

static int ap_cmp(const void* a, const void* b) {
  const uint64_t x = ((const ap_edge*)a)->key, y = ((const ap_edge*)b)->key;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static double ap_dot(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static double ap_half_angle(const double* x1, const double* v0, const double* v1) {
  const double l0 = sqrt(ap_dot(v0, v0)), l1 = sqrt(ap_dot(v1, v1));
  const double c[3] = {v0[1] * v1[2] - v0[2] * v1[1], v0[2] * v1[0] - v0[0] * v1[2], v0[0] * v1[1] - v0[1] * v1[0]};
  const double num = ap_dot(x1, c);
  const double den = ((l0 * l1 + l1 * ap_dot(x1, v0)) + l0 * ap_dot(x1, v1)) + ap_dot(v0, v1);
  return atan2(num, den);
}

typedef struct {
  const double* V;
  const int32_t* F;
  int64_t nf;
  const double* seg; /* nseg x 7 */
  int64_t nseg;
  double d[3];
  int kx, ky, kz;
  double sx, sy, sz;
  const double* O;
  double* W;
} ap_ctx;

static int ap_cross(const ap_ctx* x, const double* o, const double* A, const double* B, const double* C) {
  const double a[3] = {A[0] - o[0], A[1] - o[1], A[2] - o[2]};
  const double b[3] = {B[0] - o[0], B[1] - o[1], B[2] - o[2]};
  const double c[3] = {C[0] - o[0], C[1] - o[1], C[2] - o[2]};
  const double ax = a[x->kx] - x->sx * a[x->kz], ay = a[x->ky] - x->sy * a[x->kz];
  const double bx = b[x->kx] - x->sx * b[x->kz], by = b[x->ky] - x->sy * b[x->kz];
  const double cx = c[x->kx] - x->sx * c[x->kz], cy = c[x->ky] - x->sy * c[x->kz];
  const double U = cx * by - cy * bx, V = ax * cy - ay * cx, W = bx * ay - by * ax;
  if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return 0;
  const double det = U + V + W;
  if (det == 0.0) return 0;
  const double T = U * (x->sz * a[x->kz]) + V * (x->sz * b[x->kz]) + W * (x->sz * c[x->kz]);
  if (det > 0.0 ? T < 0.0 : T > 0.0) return 0;
  const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
  const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  return (x->d[0] * n[0] + x->d[1] * n[1]) + x->d[2] * n[2] > 0.0 ? 1 : -1;
}

static void ap_range(int64_t b, int64_t e, void* vctx) {
  const ap_ctx* x = (const ap_ctx*)vctx;
  const double x1[3] = {-x->d[0], -x->d[1], -x->d[2]};
  for (int64_t q = b; q < e; ++q) {
    const double* p = x->O + 3 * q;
    double area = 0.0;
    for (int64_t s = 0; s < x->nseg; ++s) {
      const double* g = x->seg + 7 * s;
      const double v0[3] = {g[0] - p[0], g[1] - p[1], g[2] - p[2]};
      const double v1[3] = {g[3] - p[0], g[4] - p[1], g[5] - p[2]};
      area += g[6] * ap_half_angle(x1, v0, v1);
    }
    const double frac = area / (2.0 * 3.1415926535897932384626433832795);
    int c = 0;
    for (int64_t t = 0; t < x->nf; ++t) {
      const int32_t* f = x->F + 3 * t;
      c += ap_cross(x, p, x->V + 3 * (int64_t)f[0], x->V + 3 * (int64_t)f[1], x->V + 3 * (int64_t)f[2]);
    }
    x->W[q] = frac + (double)c;
  }
}

int oracle_winding_number_antipodal(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* x0,
                                    const double* O, int64_t no, double* W, int64_t* nseg_out, int nthreads) {
  (void)nv;
  ap_ctx x;
  memset(&x, 0, sizeof x);
  if (x0) {
    for (int k = 0; k < 3; ++k) x.d[k] = x0[k];
  } else {
    const double d[3] = {1.0, sqrt(2.0), sqrt(3.0)};
    const double nrm = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) x.d[k] = d[k] / nrm;
  }
  x.kz = 0;
  for (int k = 1; k < 3; ++k)
    if (fabs(x.d[k]) > fabs(x.d[x.kz])) x.kz = k;
  x.kx = (x.kz + 1) % 3;
  x.ky = (x.kx + 1) % 3;
  if (x.d[x.kz] < 0.0) {
    const int t = x.kx;
    x.kx = x.ky;
    x.ky = t;
  }
  x.sx = x.d[x.kx] / x.d[x.kz];
  x.sy = x.d[x.ky] / x.d[x.kz];
  x.sz = 1.0 / x.d[x.kz];
  /* boundary segments */
  ap_edge* ed = (ap_edge*)malloc(sizeof(ap_edge) * (size_t)(3 * nf + 1));
  double* seg = (double*)malloc(sizeof(double) * 7 * (size_t)(3 * nf + 1));
  if (!ed || !seg) {
    free(ed);
    free(seg);
    return -2;
  }
  for (int64_t t = 0; t < nf; ++t)
    for (int e = 0; e < 3; ++e) {
      const uint32_t u = (uint32_t)F[3 * t + e], v = (uint32_t)F[3 * t + (e + 1) % 3];
      const uint32_t lo = u < v ? u : v, hi = u < v ? v : u;
      ed[3 * t + e].key = ((uint64_t)lo << 32) | hi;
      ed[3 * t + e].sgn = u < v ? 1 : -1;
    }
  qsort(ed, (size_t)(3 * nf), sizeof(ap_edge), ap_cmp);
  int64_t nseg = 0;
  for (int64_t s = 0; s < 3 * nf;) {
    int64_t t = s;
    int sum = 0;
    while (t < 3 * nf && ed[t].key == ed[s].key) sum += ed[t++].sgn;
    if (sum != 0) {
      const int64_t lo = (int64_t)(ed[s].key >> 32), hi = (int64_t)(ed[s].key & 0xFFFFFFFFull);
      const int64_t ia = sum > 0 ? lo : hi, ib = sum > 0 ? hi : lo;
      double* g = seg + 7 * nseg++;
      for (int k = 0; k < 3; ++k) {
        g[k] = V[3 * ia + k];
        g[3 + k] = V[3 * ib + k];
      }
      g[6] = (double)(sum > 0 ? sum : -sum);
    }
    s = t;
  }
  free(ed);
  x.V = V;
  x.F = F;
  x.nf = nf;
  x.seg = seg;
  x.nseg = nseg;
  x.O = O;
  x.W = W;
  oracle_parallel_for(no, ap_range, &x, 1, nthreads);
  free(seg);
  if (nseg_out) *nseg_out = nseg;
  return 0;
}
