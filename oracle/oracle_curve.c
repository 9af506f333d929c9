/*
 * oracle_curve.c -- TEST INFRASTRUCTURE (the checker, never the product).
 *
 * CPU restatement of the 2D fast winding number of a curve (a set of
 * segments), UT_SubtendedAngle<float,float>:
 *   UT_SubtendedAngle::init        FastWindingNumberForSoups.h:7342-7604
 *     (segment boxes 7365-7390, BVH<4>::init<BOX_AREA,S,2> 7392, the
 *     precompute functors item 7453-7516 / post 7518-7599)
 *   UT_SubtendedAngle::computeAngle 7700-7828 (pre 7725-7800, item 7801-7809,
 *     post 7810-7817) over BVH::traverseVector 4804-4848
 *   UTsignedAngleSegment           6279-6307
 *
 * Same operation order as the reference; built with -ffp-contract=off it
 * reproduces the compiled reference bit for bit (node arrays, BoxData, angle
 * per query), which tests/test_oracle_curve.py checks.  The 4-wide SSE lanes
 * (v4uf) are scalar loops over the 4 child slots.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define CBVH_N 4
#define CNODE_EMPTY 0xFFFFFFFFu
#define CNODE_INTERNAL 0x80000000u

/* BoxData (7287-7320), member order of the reference */
typedef struct {
  float maxP2[4];
  float P[2][4];
  float N[2][4];
  float NijDiag[2][4];
  float Nxy_Nyx[4];
  float NijkDiag[2][4];
  float N2xxy_yxx[4];
  float N2yyx_xyy[4];
} cboxdata_t;

/* LocalData (7409-7427) */
typedef struct {
  float box[2][2]; /* [axis][min, max] */
  float P[2], LP[2], N[2];
  float length;
  float NijDiag[2];
  float Nxy, Nyx;
  float NijkDiag[2];
  float N2xxy_yxx, N2yyx_xyy;
} clocal_t;

struct oracle_curve {
  int32_t np, ne;
  int order;
  float* P;         /* np x 2 */
  int32_t* E;       /* ne x 2 */
  float* seg_boxes; /* ne x 6 (z = 0) */
  uint32_t* nodes;  /* nnodes x 4 */
  uint32_t nnodes;
  cboxdata_t* data;
};

static inline float cmax(float a, float b) { return (a < b) ? b : a; }
static inline float cmin(float a, float b) { return (a > b) ? b : a; }

/* PrecomputeFunctors::item (7453-7516) */
static void cp_item(const oracle_curve* c, uint32_t itemi, clocal_t* d) {
  const float* a = c->P + 2 * (int64_t)c->E[2 * itemi];
  const float* b = c->P + 2 * (int64_t)c->E[2 * itemi + 1];
  const float ab[2] = {b[0] - a[0], b[1] - a[1]};
  const float* sb = c->seg_boxes + 6 * (size_t)itemi;
  d->box[0][0] = sb[0];
  d->box[0][1] = sb[1];
  d->box[1][0] = sb[2];
  d->box[1][1] = sb[3];
  const float N[2] = {ab[1], -ab[0]};
  const float length2 = ab[0] * ab[0] + ab[1] * ab[1];
  const float length = sqrtf(length2);
  const float P[2] = {0.5f * (a[0] + b[0]), 0.5f * (a[1] + b[1])};
  for (int k = 0; k < 2; ++k) {
    d->P[k] = P[k];
    d->LP[k] = P[k] * length;
    d->N[k] = N[k];
  }
  d->length = length;
  if (c->order < 1) return;
  d->NijDiag[0] = d->NijDiag[1] = 0.0f;
  d->Nxy = d->Nyx = 0.0f;
  if (c->order < 2) return;
  if (length == 0.0f) {
    d->NijkDiag[0] = d->NijkDiag[1] = 0.0f;
    d->N2xxy_yxx = d->N2yyx_xyy = 0.0f;
    return;
  }
  const float ixx = ab[0] * ab[0] / 12.0f;
  const float ixy = ab[0] * ab[1] / 12.0f;
  const float iyy = ab[1] * ab[1] / 12.0f;
  d->NijkDiag[0] = ixx * N[0];
  d->NijkDiag[1] = iyy * N[1];
  const float Nxxy = N[0] * ixy, Nyxx = N[1] * ixx, Nyyx = N[1] * ixy, Nxyy = N[0] * iyy;
  d->N2xxy_yxx = 2.0f * Nxxy + Nyxx;
  d->N2yyx_xyy = 2.0f * Nyyx + Nxyy;
}

/* PrecomputeFunctors::post (7518-7599) */
static void cp_post(const oracle_curve* c, uint32_t nodei, clocal_t* out, int nch, const clocal_t* ch) {
  cboxdata_t* bd = &c->data[nodei];
  float N[2] = {ch[0].N[0], ch[0].N[1]};
  float LP[2] = {ch[0].LP[0], ch[0].LP[1]};
  float length = ch[0].length;
  for (int k = 0; k < 2; ++k) {
    bd->N[k][0] = N[k];
    bd->P[k][0] = ch[0].P[k];
  }
  for (int i = 1; i < nch; ++i) {
    for (int k = 0; k < 2; ++k) {
      N[k] += ch[i].N[k];
      bd->N[k][i] = ch[i].N[k];
      LP[k] += ch[i].LP[k];
    }
    length += ch[i].length;
    for (int k = 0; k < 2; ++k) bd->P[k][i] = ch[i].P[k];
  }
  for (int i = nch; i < CBVH_N; ++i)
    for (int k = 0; k < 2; ++k) bd->N[k][i] = bd->P[k][i] = 0.0f;
  for (int k = 0; k < 2; ++k) {
    out->N[k] = N[k];
    out->LP[k] = LP[k];
  }
  out->length = length;
  float box[2][2] = {{ch[0].box[0][0], ch[0].box[0][1]}, {ch[0].box[1][0], ch[0].box[1][1]}};
  for (int i = 1; i < nch; ++i)
    for (int a = 0; a < 2; ++a) {
      box[a][0] = (box[a][0] < ch[i].box[a][0]) ? box[a][0] : ch[i].box[a][0];
      box[a][1] = (box[a][1] > ch[i].box[a][1]) ? box[a][1] : ch[i].box[a][1];
    }
  float avg[2];
  if (length > 0.0f) {
    const float r = 1.0f / length; /* UT_FixedVector::operator/ multiplies by the reciprocal (3437-3450) */
    avg[0] = LP[0] * r;
    avg[1] = LP[1] * r;
  } else {
    avg[0] = 0.5f * (box[0][0] + box[0][1]);
    avg[1] = 0.5f * (box[1][0] + box[1][1]);
  }
  out->P[0] = avg[0];
  out->P[1] = avg[1];
  memcpy(out->box, box, sizeof(box));
  for (int i = 0; i < nch; ++i) {
    const float d0 = cmax(ch[i].P[0] - ch[i].box[0][0], ch[i].box[0][1] - ch[i].P[0]);
    const float d1 = cmax(ch[i].P[1] - ch[i].box[1][0], ch[i].box[1][1] - ch[i].P[1]);
    bd->maxP2[i] = d0 * d0 + d1 * d1;
  }
  for (int i = nch; i < CBVH_N; ++i) bd->maxP2[i] = INFINITY;
  if (c->order < 1) return;
  out->NijDiag[0] = ch[0].NijDiag[0];
  out->NijDiag[1] = ch[0].NijDiag[1];
  out->Nxy = out->Nyx = 0.0f;
  out->NijkDiag[0] = ch[0].NijkDiag[0];
  out->NijkDiag[1] = ch[0].NijkDiag[1];
  out->N2xxy_yxx = ch[0].N2xxy_yxx;
  out->N2yyx_xyy = ch[0].N2yyx_xyy;
  for (int i = 1; i < nch; ++i) {
    for (int k = 0; k < 2; ++k) {
      out->NijDiag[k] += ch[i].NijDiag[k];
      out->NijkDiag[k] += ch[i].NijkDiag[k];
    }
    out->N2xxy_yxx += ch[i].N2xxy_yxx;
    out->N2yyx_xyy += ch[i].N2yyx_xyy;
  }
  for (int i = 0; i < CBVH_N; ++i) {
    const int on = i < nch;
    for (int k = 0; k < 2; ++k) {
      bd->NijDiag[k][i] = on ? ch[i].NijDiag[k] : 0.0f;
      bd->NijkDiag[k][i] = on ? ch[i].NijkDiag[k] : 0.0f;
    }
    bd->Nxy_Nyx[i] = on ? ch[i].Nxy + ch[i].Nyx : 0.0f;
    bd->N2xxy_yxx[i] = on ? ch[i].N2xxy_yxx : 0.0f;
    bd->N2yyx_xyy[i] = on ? ch[i].N2yyx_xyy : 0.0f;
  }
  for (int i = 0; i < nch; ++i) {
    const clocal_t* d = &ch[i];
    const float dp[2] = {d->P[0] - out->P[0], d->P[1] - out->P[1]};
    const float* n = d->N;
    out->NijDiag[0] += n[0] * dp[0];
    out->NijDiag[1] += n[1] * dp[1];
    const float Nxy = d->Nxy + n[0] * dp[1];
    const float Nyx = d->Nyx + n[1] * dp[0];
    out->Nxy += Nxy;
    out->Nyx += Nyx;
    if (c->order >= 2) {
      for (int k = 0; k < 2; ++k) out->NijkDiag[k] += (2.0f * dp[k]) * d->NijDiag[k] + (dp[k] * dp[k]) * n[k];
      out->N2xxy_yxx += 2.0f * (dp[1] * d->NijDiag[0] + dp[0] * d->Nxy + n[0] * dp[0] * dp[1]) +
                        2.0f * d->Nyx * dp[0] + n[1] * dp[0] * dp[0];
      out->N2yyx_xyy += 2.0f * (dp[0] * d->NijDiag[1] + dp[1] * d->Nyx + n[1] * dp[1] * dp[0]) +
                        2.0f * d->Nxy * dp[1] + n[0] * dp[1] * dp[1];
    }
  }
}

/* BVH::traverseHelper (4641-4667) with the precompute functors */
static void cp_traverse(const oracle_curve* c, uint32_t nodei, clocal_t* out) {
  clocal_t local[CBVH_N];
  memset(local, 0, sizeof(local));
  const uint32_t* node = c->nodes + 4 * (size_t)nodei;
  int sidx;
  for (sidx = 0; sidx < CBVH_N; ++sidx) {
    const uint32_t ni = node[sidx];
    if (ni & CNODE_INTERNAL) {
      if (ni == CNODE_EMPTY) break;
      cp_traverse(c, ni & ~CNODE_INTERNAL, &local[sidx]);
    } else {
      cp_item(c, ni, &local[sidx]);
    }
  }
  cp_post(c, nodei, out, sidx, local);
}

oracle_curve* oracle_curve_build(const float* P, int32_t np, const int32_t* E, int32_t ne, int order) {
  oracle_curve* c = (oracle_curve*)calloc(1, sizeof(oracle_curve));
  c->np = np;
  c->ne = ne;
  c->order = order;
  c->P = (float*)malloc(sizeof(float) * 2 * (size_t)(np > 0 ? np : 1));
  c->E = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)(ne > 0 ? ne : 1));
  memcpy(c->P, P, sizeof(float) * 2 * (size_t)np);
  memcpy(c->E, E, sizeof(int32_t) * 2 * (size_t)ne);
  /* segment boxes (7365-7390): initBounds(p0), enlargeBounds(p1) */
  c->seg_boxes = (float*)calloc(6 * (size_t)(ne > 0 ? ne : 1), sizeof(float));
  for (int32_t i = 0; i < ne; ++i) {
    const float* p0 = c->P + 2 * (int64_t)c->E[2 * i];
    const float* p1 = c->P + 2 * (int64_t)c->E[2 * i + 1];
    float* b = c->seg_boxes + 6 * (size_t)i;
    for (int a = 0; a < 2; ++a) {
      b[2 * a] = cmin(p0[a], p1[a]);
      b[2 * a + 1] = cmax(p0[a], p1[a]);
    }
  }
  c->nnodes = oracle_bvh4_build(c->seg_boxes, (uint32_t)(ne > 0 ? ne : 0), 2, &c->nodes);
  c->data = (cboxdata_t*)calloc(c->nnodes > 0 ? c->nnodes : 1, sizeof(cboxdata_t));
  if (c->nnodes > 0) {
    clocal_t root;
    memset(&root, 0, sizeof(root));
    cp_traverse(c, 0, &root);
  }
  return c;
}

void oracle_curve_free(oracle_curve* c) {
  if (!c) return;
  free(c->P);
  free(c->E);
  free(c->seg_boxes);
  free(c->nodes);
  free(c->data);
  free(c);
}

int32_t oracle_curve_num_nodes(const oracle_curve* c) { return (int32_t)c->nnodes; }
void oracle_curve_nodes(const oracle_curve* c, uint32_t* out) {
  if (c->nnodes) memcpy(out, c->nodes, sizeof(uint32_t) * 4 * (size_t)c->nnodes);
}
void oracle_curve_boxdata(const oracle_curve* c, float* out) {
  if (c->nnodes) memcpy(out, c->data, sizeof(cboxdata_t) * (size_t)c->nnodes);
}

/* UTsignedAngleSegment (6279-6307) in float */
static float angle_segment(const float* a, const float* b, const float* q) {
  const float qa[2] = {a[0] - q[0], a[1] - q[1]};
  const float qb[2] = {b[0] - q[0], b[1] - q[1]};
  if ((qa[0] == 0.0f && qa[1] == 0.0f) || (qb[0] == 0.0f && qb[1] == 0.0f)) return 0.0f;
  const float num = qa[0] * qb[1] - qa[1] * qb[0];
  if (num == 0.0f) return 0.0f;
  const float den = qa[0] * qb[0] + qa[1] * qb[1];
  return atan2f(num, den);
}

typedef struct {
  const oracle_curve* c;
  float q[2];
  float acc2;
  int64_t visits, far_accepted, exact;
} ca_ctx;

/* AngleFunctors::pre (7725-7800) */
static uint32_t ca_pre(ca_ctx* x, uint32_t nodei, float* data_for_parent) {
  const cboxdata_t* d = &x->c->data[nodei];
  x->visits++;
  float q[2][4], ql2[4];
  uint32_t descend = 0;
  for (int l = 0; l < 4; ++l) {
    q[0][l] = x->q[0] - d->P[0][l];
    q[1][l] = x->q[1] - d->P[1][l];
    ql2[l] = q[0][l] * q[0][l] + q[1][l] * q[1][l];
    if (ql2[l] <= d->maxP2[l] * x->acc2) descend |= 1u << l;
  }
  if (descend == 0xFu) {
    *data_for_parent = 0.0f;
    return 0xFu;
  }
  float om[4];
  for (int l = 0; l < 4; ++l) {
    const float m2 = 1.0f / ql2[l];
    const float m1 = sqrtf(m2);
    const float qx = q[0][l] * m1, qy = q[1][l] * m1;
    float o = -m1 * (qx * d->N[0][l] + qy * d->N[1][l]);
    if (x->c->order >= 1) {
      const float q2x = qx * qx, q2y = qy * qy;
      const float o1 = m2 * ((d->NijDiag[0][l] + d->NijDiag[1][l]) -
                             2.0f * ((q2x * d->NijDiag[0][l] + q2y * d->NijDiag[1][l]) + qx * qy * d->Nxy_Nyx[l]));
      o += o1;
      if (x->c->order >= 2) {
        const float q3x = q2x * qx, q3y = q2y * qy;
        const float m3 = m2 * m1;
        const float t0x = d->N2yyx_xyy[l], t0y = d->N2xxy_yxx[l];
        const float t1x = qy * d->N2xxy_yxx[l], t1y = qx * d->N2yyx_xyy[l];
        const float sx = 3.0f * d->NijkDiag[0][l] + t0x, sy = 3.0f * d->NijkDiag[1][l] + t0y;
        const float o2 = m3 * ((qx * sx + qy * sy) -
                               4.0f * ((q3x * d->NijkDiag[0][l] + q3y * d->NijkDiag[1][l]) + (q2x * t1x + q2y * t1y)));
        o += o2;
      }
    }
    om[l] = o;
  }
  uint32_t mask = 0;
  for (int l = 0; l < 4; ++l) {
    if (isfinite(om[l]) && !((descend >> l) & 1u)) mask |= 1u << l;
    if (!((mask >> l) & 1u)) om[l] = 0.0f;
  }
  for (int l = 0; l < 4; ++l)
    if ((mask >> l) & 1u) x->far_accepted++;
  float sum = om[0];
  for (int l = 1; l < 4; ++l) sum += om[l];
  *data_for_parent = sum;
  return (~mask) & 0xFu;
}

/* BVH::traverseVectorHelper (4818-4848) with the angle functors' item/post (7801-7817) */
static void ca_traverse(ca_ctx* x, uint32_t nodei, float* data_for_parent) {
  const uint32_t* node = x->c->nodes + 4 * (size_t)nodei;
  uint32_t descend = ca_pre(x, nodei, data_for_parent);
  if (!descend) return;
  float local[CBVH_N];
  int sidx;
  for (sidx = 0; sidx < CBVH_N; ++sidx) {
    if ((descend >> sidx) & 1u) {
      const uint32_t ni = node[sidx];
      if (ni & CNODE_INTERNAL) {
        if (ni == CNODE_EMPTY) {
          descend &= (1u << sidx) - 1;
          break;
        }
        ca_traverse(x, ni & ~CNODE_INTERNAL, &local[sidx]);
      } else {
        x->exact++;
        local[sidx] = angle_segment(x->c->P + 2 * (int64_t)x->c->E[2 * ni], x->c->P + 2 * (int64_t)x->c->E[2 * ni + 1],
                                    x->q);
      }
    }
  }
  float sum = (descend & 1u) ? local[0] : 0.0f;
  for (int i = 1; i < sidx; ++i) sum += ((descend >> i) & 1u) ? local[i] : 0.0f;
  *data_for_parent += sum;
}

static float curve_angle_one(const oracle_curve* c, const float* q, float acc, ca_ctx* x) {
  x->c = c;
  x->q[0] = q[0];
  x->q[1] = q[1];
  x->acc2 = acc * acc;
  x->visits = x->far_accepted = x->exact = 0;
  /* an empty tree: the reference returns an uninitialised value; 0 here */
  if (c->nnodes == 0) return 0.0f;
  float sum;
  ca_traverse(x, 0, &sum);
  return sum;
}

typedef struct {
  const oracle_curve* c;
  const float* Q;
  float acc;
  float* out;
} cq_ctx;

static void cq_range(int64_t b, int64_t e, void* vctx) {
  cq_ctx* cx = (cq_ctx*)vctx;
  ca_ctx x;
  for (int64_t p = b; p < e; ++p) cx->out[p] = curve_angle_one(cx->c, cx->Q + 2 * p, cx->acc, &x);
}

void oracle_curve_angle(const oracle_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle,
                        int nthreads) {
  cq_ctx cx = {c, Q, accuracy_scale, angle};
  oracle_parallel_for(nq, cq_range, &cx, 1000, nthreads);
}

void oracle_curve_count(const oracle_curve* c, const float* q, float accuracy_scale, int64_t* visits,
                        int64_t* far_accepted, int64_t* exact_segments) {
  ca_ctx x;
  curve_angle_one(c, q, accuracy_scale, &x);
  *visits = x.visits;
  *far_accepted = x.far_accepted;
  *exact_segments = x.exact;
}
