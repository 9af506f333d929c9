/*
 * oracle_common.c -- TEST INFRASTRUCTURE (the checker, never the product).
 *
 * Restates igl::parallel_for (parallel_for.h:140-204) and
 * igl::default_num_threads (default_num_threads.cpp:45-59), and the exact
 * FP64 solid angle / brute-force winding number
 * (solid_angle.cpp:10-55, winding_number.cpp:62-89).
 */
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>

/* default_num_threads.cpp:45-59: IGL_NUM_THREADS, else hardware concurrency, else 8 */
int oracle_default_num_threads(void) {
  const char* s = getenv("IGL_NUM_THREADS");
  if (s) {
    int v = atoi(s);
    if (v > 0) return v;
  }
  long hw = sysconf(_SC_NPROCESSORS_ONLN);
  return hw > 0 ? (int)hw : 8;
}

typedef struct {
  oracle_range_fn fn;
  void* ctx;
  int64_t b, e;
} pf_job;

static void* pf_run(void* arg) {
  pf_job* j = (pf_job*)arg;
  j->fn(j->b, j->e, j->ctx);
  return NULL;
}

/* parallel_for.h:140-204.  Slice = max(round((n+1)/T), 1); threads t < T-1 take
 * consecutive slices while work remains; the last thread takes the remainder. */
int oracle_parallel_for(int64_t n, oracle_range_fn fn, void* ctx, int64_t min_parallel, int nthreads) {
  if (n <= 0) return 0;
  int T = nthreads > 0 ? nthreads : oracle_default_num_threads();
  if (n < min_parallel || T <= 1) {
    fn(0, n, ctx);
    return 1;
  }
  int64_t slice = (int64_t)llround((double)(n + 1) / (double)T);
  if (slice < 1) slice = 1;
  pf_job* jobs = (pf_job*)malloc(sizeof(pf_job) * (size_t)T);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)T);
  int used = 0;
  int64_t i1 = 0, i2 = slice < n ? slice : n;
  int t = 0;
  for (; t + 1 < T && i1 < n; ++t) {
    jobs[used] = (pf_job){fn, ctx, i1, i2};
    pthread_create(&th[used], NULL, pf_run, &jobs[used]);
    ++used;
    i1 = i2;
    i2 = (i2 + slice) < n ? (i2 + slice) : n;
  }
  if (i1 < n) {
    jobs[used] = (pf_job){fn, ctx, i1, n};
    pthread_create(&th[used], NULL, pf_run, &jobs[used]);
    ++used;
  }
  for (int k = 0; k < used; ++k) pthread_join(th[k], NULL);
  free(jobs);
  free(th);
  return used;
}

/* igl::PI (PI.h) */
#define ORACLE_PI 3.1415926535897932384626433832795

/* solid_angle.cpp:10-55.  v rows are the corners relative to P; vl their
 * norms; det by the rule of Sarrus in the reference's term order; result is
 * atan2(det, |a||b||c| + (b.c)|a| + (c.a)|b| + (a.b)|c|) / (2 pi). */
double oracle_solid_angle(const double* A, const double* B, const double* C, const double* P) {
  double v[3][3];
  for (int d = 0; d < 3; ++d) {
    v[0][d] = A[d] - P[d];
    v[1][d] = B[d] - P[d];
    v[2][d] = C[d] - P[d];
  }
  double vl[3];
  for (int r = 0; r < 3; ++r) vl[r] = sqrt(v[r][0] * v[r][0] + v[r][1] * v[r][1] + v[r][2] * v[r][2]);
  double detf = v[0][0] * v[1][1] * v[2][2] + v[1][0] * v[2][1] * v[0][2] + v[2][0] * v[0][1] * v[1][2] -
                v[2][0] * v[1][1] * v[0][2] - v[1][0] * v[0][1] * v[2][2] - v[0][0] * v[2][1] * v[1][2];
  double dp[3];
  dp[0] = v[1][0] * v[2][0];
  dp[0] += v[1][1] * v[2][1];
  dp[0] += v[1][2] * v[2][2];
  dp[1] = v[2][0] * v[0][0];
  dp[1] += v[2][1] * v[0][1];
  dp[1] += v[2][2] * v[0][2];
  dp[2] = v[0][0] * v[1][0];
  dp[2] += v[0][1] * v[1][1];
  dp[2] += v[0][2] * v[1][2];
  return atan2(detf, vl[0] * vl[1] * vl[2] + dp[0] * vl[0] + dp[1] * vl[1] + dp[2] * vl[2]) / (2. * ORACLE_PI);
}

/* UTsignedSolidAngleTri (FastWindingNumberForSoups.h:6070-6110) in double:
 * normalise by multiplying with the reciprocal length (UT_FixedVector
 * operator/=, 3423-3435), stable numerator qa.((qb-qa)x(qc-qa)). */
double oracle_solid_angle_tri_f64(const double* a, const double* b, const double* c, const double* q) {
  double qa[3], qb[3], qc[3];
  for (int i = 0; i < 3; ++i) {
    qa[i] = a[i] - q[i];
    qb[i] = b[i] - q[i];
    qc[i] = c[i] - q[i];
  }
  double al = sqrt(qa[0] * qa[0] + qa[1] * qa[1] + qa[2] * qa[2]);
  double bl = sqrt(qb[0] * qb[0] + qb[1] * qb[1] + qb[2] * qb[2]);
  double cl = sqrt(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2]);
  if (al == 0 || bl == 0 || cl == 0) return 0.0;
  double ra = 1 / al, rb = 1 / bl, rc = 1 / cl;
  for (int i = 0; i < 3; ++i) {
    qa[i] *= ra;
    qb[i] *= rb;
    qc[i] *= rc;
  }
  double u[3] = {qb[0] - qa[0], qb[1] - qa[1], qb[2] - qa[2]};
  double w[3] = {qc[0] - qa[0], qc[1] - qa[1], qc[2] - qa[2]};
  double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
  double num = qa[0] * cr[0] + qa[1] * cr[1] + qa[2] * cr[2];
  if (num == 0) return 0.0;
  double den = 1.0 + (qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2]) + (qa[0] * qc[0] + qa[1] * qc[1] + qa[2] * qc[2]) +
               (qb[0] * qc[0] + qb[1] * qc[1] + qb[2] * qc[2]);
  return 2.0 * atan2(num, den);
}

typedef struct {
  const double* V;
  const int32_t* F;
  int64_t nf;
  const double* O;
  double* W;
} bf_ctx;

static void bf_range(int64_t b, int64_t e, void* vctx) {
  bf_ctx* c = (bf_ctx*)vctx;
  for (int64_t o = b; o < e; ++o) {
    const double* p = c->O + 3 * o;
    double w = 0;
    for (int64_t f = 0; f < c->nf; ++f) {
      const int32_t* t = c->F + 3 * f;
      w += oracle_solid_angle(c->V + 3 * (int64_t)t[0], c->V + 3 * (int64_t)t[1], c->V + 3 * (int64_t)t[2], p);
    }
    c->W[o] = w;
  }
}

/* winding_number.cpp:62-89 (single point, sum over F rows in order), batched
 * with parallel_for(..., 10000) as winding_number.cpp:48-51 batches. */
void oracle_winding_number_bf(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                              int64_t no, double* W, int nthreads) {
  (void)nv;
  bf_ctx c = {V, F, nf, O, W};
  oracle_parallel_for(no, bf_range, &c, 10000, nthreads);
}
