/*
 * oracle_soup.c -- TEST INFRASTRUCTURE (the checker, never the product).
 *
 * CPU restatement of the triangle-soup fast winding number:
 *   UT::BVH<4>::init<BOX_AREA>   FastWindingNumberForSoups.h:4579-5944
 *   UT_SolidAngle::init          FastWindingNumberForSoups.h:6543-7133
 *   UT_SolidAngle::computeSolidAngle + BVH::traverseVector  7148-7284, 4804-4848
 *   UTsignedSolidAngleTri        6070-6110
 *   fast_winding_number(bvh, accuracy_scale, Q, W)  fast_winding_number.cpp:418-437
 *
 * The restatement follows the reference's operation order so that, built with
 * -ffp-contract=off, it reproduces the compiled reference bit for bit (node
 * arrays, BoxData and per-query Omega); tests/test_oracle.py checks this
 * (test_soup_oracle_bit_exact_vs_reference and the two tests after it).
 * The 4-wide SSE lanes of the reference (v4uf) are restated as scalar loops
 * over the 4 child slots: each SSE lane operation is an independent IEEE
 * single-precision operation, so the result is identical.
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/sysinfo.h>

#define BVH_N 4
#define NODE_EMPTY 0xFFFFFFFFu
#define NODE_INTERNAL 0x80000000u
#define NSPANS 16
#define NSPLITS 15
#define MIN_FRACTION 16
#define SMALL_LIMIT 6
#define MID_LIMIT 32

typedef struct {
  float v[3][2]; /* [axis][min,max], UT::Box<float,3> (3914-4162) */
} obox;

typedef struct {
  uint32_t child[BVH_N];
} onode;

/* h_min / h_max (250-256): on equality return the first argument */
static inline float hmin(float a, float b) { return (a > b) ? b : a; }
static inline float hmax(float a, float b) { return (a < b) ? b : a; }

static inline void box_init_empty(obox* b) {
  for (int a = 0; a < 3; ++a) {
    b->v[a][0] = FLT_MAX;
    b->v[a][1] = -FLT_MAX;
  }
}
/* Box::combine (3983-3992) */
static inline void box_combine(obox* b, const obox* s) {
  for (int a = 0; a < 3; ++a) {
    float mn = b->v[a][0], mx = b->v[a][1];
    b->v[a][0] = (mn < s->v[a][0]) ? mn : s->v[a][0];
    b->v[a][1] = (mx > s->v[a][1]) ? mx : s->v[a][1];
  }
}
/* Box::half_surface_area, NAXES==2 (4071-4075) and NAXES==3 (4076-4080) */
static inline float box_area(const obox* b, int naxes) {
  const float d0 = b->v[0][1] - b->v[0][0];
  const float d1 = b->v[1][1] - b->v[1][0];
  if (naxes == 2) return d0 + d1;
  const float d2 = b->v[2][1] - b->v[2][0];
  return d0 * d1 + d1 * d2 + d2 * d0;
}
/* utBoxCenter for a Box: min+max, i.e. 2x the centre (4500-4504) */
static inline float box_cx2(const obox* b, uint32_t axis) { return b->v[axis][0] + b->v[axis][1]; }

/* utBoxExclude for float boxes (4489-4499): any exponent field all ones */
static inline int box_exclude(const obox* b) {
  const uint32_t* p = (const uint32_t*)b;
  int r = 0;
  for (int i = 0; i < 6; ++i) r |= ((p[i] & 0x7F800000u) == 0x7F800000u);
  return r;
}

/* ---------------- growable node array ---------------- */
typedef struct {
  onode* a;
  uint32_t n, cap;
} nodevec;

static uint32_t nv_push(nodevec* v) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 64;
    v->a = (onode*)realloc(v->a, sizeof(onode) * v->cap);
  }
  return v->n++;
}

/* ---------------- partitionByCentre (5900-5944) ---------------- */
static void partition_by_centre(const obox* boxes, uint32_t* indices, uint32_t* indices_end, uint32_t axis,
                                float pivotx2, uint32_t** pstart, uint32_t** pend) {
  uint32_t* pivot_start = indices;
  uint32_t* pivot_end = indices;
  for (uint32_t* ps = indices; ps != indices_end; ++ps) {
    const float s = box_cx2(&boxes[*ps], axis);
    if (s < pivotx2) {
      if (ps != pivot_start) {
        if (pivot_start == pivot_end) {
          const uint32_t t = *ps;
          *ps = *pivot_start;
          *pivot_start = t;
        } else {
          const uint32_t t = *ps;
          *ps = *pivot_end;
          *pivot_end = *pivot_start;
          *pivot_start = t;
        }
      }
      ++pivot_start;
      ++pivot_end;
    } else if (s == pivotx2) {
      if (ps != pivot_end) {
        const uint32_t t = *ps;
        *ps = *pivot_end;
        *pivot_end = t;
      }
      ++pivot_end;
    }
  }
  *pstart = pivot_start;
  *pend = pivot_end;
}

/* ---------------- nthElement (5845-5898) ---------------- */
static void nth_element(const obox* boxes, uint32_t* indices, uint32_t* indices_end, uint32_t axis, uint32_t* nth) {
  for (;;) {
    float p[3] = {box_cx2(&boxes[indices[0]], axis), box_cx2(&boxes[indices[(indices_end - indices) / 2]], axis),
                  box_cx2(&boxes[*(indices_end - 1)], axis)};
    float t;
    if (p[0] < p[1]) { t = p[0]; p[0] = p[1]; p[1] = t; }
    if (p[0] < p[2]) { t = p[0]; p[0] = p[2]; p[2] = t; }
    if (p[1] < p[2]) { t = p[1]; p[1] = p[2]; p[2] = t; }
    uint32_t *ps, *pe;
    partition_by_centre(boxes, indices, indices_end, axis, p[1], &ps, &pe);
    if (nth < ps) {
      indices_end = ps;
    } else if (nth < pe) {
      return;
    } else {
      indices = pe;
    }
    if (indices_end <= indices + 1) return;
  }
}

/* ---------------- split (5348-5799) ---------------- */
static void split2(obox axes_minmax, const obox* boxes, uint32_t* indices, uint32_t nboxes, uint32_t** split_indices,
                   obox* split_boxes, int naxes) {
  if (nboxes == 2) {
    split_boxes[0] = boxes[indices[0]];
    split_boxes[1] = boxes[indices[1]];
    *split_indices = indices + 1;
    return;
  }
  if (nboxes <= SMALL_LIMIT) {
    /* exhaustive over the 2^(n-1)-1 partitions with box 0 fixed in part 0 (5362-5451) */
    obox local_boxes[SMALL_LIMIT];
    for (uint32_t b = 0; b < nboxes; ++b) local_boxes[b] = boxes[indices[b]];
    const uint32_t limit = 1u << (nboxes - 1);
    uint32_t best = 0xFFFFFFFFu;
    float best_h = 0;
    for (uint32_t bits = 1; bits < limit; ++bits) {
      obox sb[2];
      sb[0] = local_boxes[0];
      box_init_empty(&sb[1]);
      uint32_t cnt[2] = {1, 0};
      for (uint32_t bit = 0; bit < nboxes - 1; ++bit) {
        uint32_t dest = (bits >> bit) & 1u;
        box_combine(&sb[dest], &local_boxes[bit + 1]);
        ++cnt[dest];
      }
      const float h = box_area(&sb[0], naxes) * (float)cnt[0] + box_area(&sb[1], naxes) * (float)cnt[1];
      if (best == 0xFFFFFFFFu || h < best_h) {
        best = bits;
        best_h = h;
        split_boxes[0] = sb[0];
        split_boxes[1] = sb[1];
      }
    }
    uint32_t local_idx[SMALL_LIMIT - 1];
    for (uint32_t b = 0; b < nboxes - 1; ++b) local_idx[b] = indices[b + 1];
    uint32_t* dst = indices + 1;
    for (uint32_t bit = 0; bit < nboxes - 1; ++bit)
      if (!((best >> bit) & 1u)) *dst++ = local_idx[bit];
    *split_indices = dst;
    for (uint32_t bit = 0; bit < nboxes - 1; ++bit)
      if ((best >> bit) & 1u) *dst++ = local_idx[bit];
    return;
  }

  uint32_t max_axis = 0;
  float max_len = axes_minmax.v[0][1] - axes_minmax.v[0][0];
  for (uint32_t a = 1; a < (uint32_t)naxes; ++a) {
    const float l = axes_minmax.v[a][1] - axes_minmax.v[a][0];
    if (l > max_len) {
      max_axis = a;
      max_len = l;
    }
  }
  if (!(max_len > 0.0f)) {
    *split_indices = indices + nboxes / 2;
    split_boxes[0] = axes_minmax;
    split_boxes[1] = axes_minmax;
    return;
  }
  const uint32_t axis = max_axis;

  if (nboxes <= MID_LIMIT) {
    /* stable sort by 2x centre (the reference's chunked insertion sort +
     * std::merge is a stable sort, 5480-5535), then try every split (5537-5578) */
    float key[MID_LIMIT];
    uint32_t li[MID_LIMIT];
    for (uint32_t i = 0; i < nboxes; ++i) {
      key[i] = box_cx2(&boxes[indices[i]], axis);
      li[i] = i;
    }
    for (uint32_t i = 1; i < nboxes; ++i) {
      uint32_t x = li[i];
      int32_t j = (int32_t)i - 1;
      while (j >= 0 && key[x] < key[li[j]]) {
        li[j + 1] = li[j];
        --j;
      }
      li[j + 1] = x;
    }
    uint32_t sorted[MID_LIMIT];
    for (uint32_t i = 0; i < nboxes; ++i) sorted[i] = indices[li[i]];
    for (uint32_t i = 0; i < nboxes; ++i) indices[i] = sorted[i];

    obox left[MID_LIMIT - 1], right[MID_LIMIT - 1];
    const uint32_t nsplits = nboxes - 1;
    obox acc = boxes[sorted[0]];
    left[0] = acc;
    for (uint32_t i = 1; i < nsplits; ++i) {
      box_combine(&acc, &boxes[sorted[i]]);
      left[i] = acc;
    }
    /* Right accumulator as the reference builds it (5561-5566): it starts from
     * sorted[nsplits-1] (not sorted[nboxes-1]) and walks down to 1, so
     * right[k] = union(sorted[k+1 .. nboxes-2]) and the last sorted box is
     * never included.  Reproduced as is: these boxes become the children's
     * axes_minmax and steer their later splits. */
    acc = boxes[sorted[nsplits - 1]];
    right[nsplits - 1] = acc;
    for (uint32_t i = nsplits - 1; i > 0; --i) {
      box_combine(&acc, &boxes[sorted[i]]);
      right[i - 1] = acc;
    }
    uint32_t best_split = 0;
    float best_h = box_area(&left[0], naxes) + box_area(&right[0], naxes) * (float)(nboxes - 1);
    for (uint32_t s = 1; s < nsplits; ++s) {
      const float h = box_area(&left[s], naxes) * (float)(s + 1) + box_area(&right[s], naxes) * (float)(nboxes - (s + 1));
      if (h < best_h) {
        best_split = s;
        best_h = h;
      }
    }
    *split_indices = indices + best_split + 1;
    split_boxes[0] = left[best_split];
    split_boxes[1] = right[best_split];
    return;
  }

  /* binned SAH over 16 spans (5580-5799) */
  const float axis_min = axes_minmax.v[max_axis][0];
  const float axis_length = max_len;
  obox span_boxes[NSPANS];
  uint32_t span_counts[NSPANS];
  for (int i = 0; i < NSPANS; ++i) {
    box_init_empty(&span_boxes[i]);
    span_counts[i] = 0;
  }
  const float axis_min_x2 = 2.0f * axis_min;
  const float axis_index_scale = (0.5f * (float)NSPANS) / axis_length;
  for (uint32_t k = 0; k < nboxes; ++k) {
    const obox* b = &boxes[indices[k]];
    const float sum = box_cx2(b, axis);
    int si = (int)((sum - axis_min_x2) * axis_index_scale);
    si = (si <= 0) ? 0 : ((si >= NSPANS - 1) ? NSPANS - 1 : si);
    ++span_counts[si];
    box_combine(&span_boxes[si], b);
  }
  obox left[NSPLITS], right[NSPLITS];
  obox acc = span_boxes[0];
  left[0] = acc;
  for (int i = 1; i < NSPLITS; ++i) {
    box_combine(&acc, &span_boxes[i]);
    left[i] = acc;
  }
  acc = span_boxes[NSPANS - 1];
  right[NSPLITS - 1] = acc;
  for (int i = NSPLITS - 1; i > 0; --i) {
    box_combine(&acc, &span_boxes[i]);
    right[i - 1] = acc;
  }
  uint32_t left_counts[NSPLITS];
  uint32_t ca = span_counts[0];
  left_counts[0] = ca;
  for (int i = 1; i < NSPLITS; ++i) {
    ca += span_counts[i];
    left_counts[i] = ca;
  }
  const uint32_t min_count = nboxes / MIN_FRACTION;
  const uint32_t max_count = (uint32_t)(((uint64_t)(MIN_FRACTION - 1) * (uint64_t)nboxes) / MIN_FRACTION);
  float smallest = INFINITY;
  uint32_t split_index = 0xFFFFFFFFu;
  for (uint32_t s = 0; s < NSPLITS; ++s) {
    const uint32_t lc = left_counts[s];
    if (lc < min_count || lc > max_count) continue;
    const uint32_t rc = nboxes - lc;
    const float h = (float)lc * box_area(&left[s], naxes) + (float)rc * box_area(&right[s], naxes);
    if (h < smallest) {
      smallest = h;
      split_index = s;
    }
  }
  uint32_t* const indices_end = indices + nboxes;
  if (split_index == 0xFFFFFFFFu) {
    uint32_t* nth;
    if (left_counts[0] > max_count)
      nth = indices + max_count;
    else if (left_counts[NSPLITS - 1] < min_count)
      nth = indices + min_count;
    else
      nth = indices + nboxes / 2;
    nth_element(boxes, indices, indices_end, max_axis, nth);
    *split_indices = nth;
    obox lb = boxes[indices[0]];
    for (uint32_t* p = indices + 1; p < nth; ++p) box_combine(&lb, &boxes[*p]);
    obox rb = boxes[nth[0]];
    for (uint32_t* p = nth + 1; p < indices_end; ++p) box_combine(&rb, &boxes[*p]);
    split_boxes[0] = lb;
    split_boxes[1] = rb;
  } else {
    const float pivotx2 = axis_min_x2 + ((float)(split_index + 1) * axis_length) / (float)(NSPANS / 2);
    uint32_t *ps, *pe;
    partition_by_centre(boxes, indices, indices_end, max_axis, pivotx2, &ps, &pe);
    uint32_t* si = indices + left_counts[split_index];
    if (si >= ps && si <= pe) {
      *split_indices = si;
      split_boxes[0] = left[split_index];
      split_boxes[1] = right[split_index];
      return;
    }
    if (si < ps)
      si = ps;
    else
      si = pe;
    if (si == indices)
      ++si;
    else if (si == indices_end)
      --si;
    *split_indices = si;
    obox lb = boxes[indices[0]];
    for (uint32_t* p = indices + 1; p < si; ++p) box_combine(&lb, &boxes[*p]);
    obox rb = boxes[si[0]];
    for (uint32_t* p = si + 1; p < indices_end; ++p) box_combine(&rb, &boxes[*p]);
    split_boxes[0] = lb;
    split_boxes[1] = rb;
  }
}

/* ---------------- multiSplit, BOX_AREA branch (5248-5346) ---------------- */
static void multi_split(const obox* axes_minmax, const obox* boxes, uint32_t* indices, uint32_t nboxes,
                        uint32_t* sub_indices[BVH_N + 1], obox sub_boxes[BVH_N], int naxes) {
  sub_indices[0] = indices;
  sub_indices[2] = indices + nboxes;
  split2(*axes_minmax, boxes, indices, nboxes, &sub_indices[1], &sub_boxes[0], naxes);
  float areas[BVH_N];
  areas[0] = box_area(&sub_boxes[0], naxes);
  areas[1] = box_area(&sub_boxes[1], naxes);
  for (uint32_t nsub = 2; nsub < BVH_N; ++nsub) {
    uint32_t choice = 0xFFFFFFFFu;
    float maxh = 0;
    for (uint32_t i = 0; i < nsub; ++i) {
      const uint32_t cnt = (uint32_t)(sub_indices[i + 1] - sub_indices[i]);
      if (cnt > 1) {
        const float h = areas[i] * (float)cnt;
        if (choice == 0xFFFFFFFFu || h > maxh) {
          choice = i;
          maxh = h;
        }
      }
    }
    uint32_t* sel_start = sub_indices[choice];
    uint32_t* sel_end = sub_indices[choice + 1];
    for (uint32_t i = nsub; i > choice; --i) sub_indices[i + 1] = sub_indices[i];
    for (uint32_t i = nsub - 1; i > choice; --i) sub_boxes[i + 1] = sub_boxes[i];
    for (uint32_t i = nsub - 1; i > choice; --i) areas[i + 1] = areas[i];
    obox parent = sub_boxes[choice];
    split2(parent, boxes, sel_start, (uint32_t)(sel_end - sel_start), &sub_indices[choice + 1], &sub_boxes[choice],
           naxes);
    areas[choice] = box_area(&sub_boxes[choice], naxes);
    areas[choice + 1] = box_area(&sub_boxes[choice + 1], naxes);
  }
}

/* ---------------- initNode (4916-5047), serial order ---------------- */
static void init_node(nodevec* nodes, uint32_t nodei, const obox* axes_minmax, const obox* boxes, uint32_t* indices,
                      uint32_t nboxes, int naxes) {
  if (nboxes <= BVH_N) {
    for (uint32_t i = 0; i < nboxes; ++i) nodes->a[nodei].child[i] = indices[i];
    for (uint32_t i = nboxes; i < BVH_N; ++i) nodes->a[nodei].child[i] = NODE_EMPTY;
    return;
  }
  uint32_t* sub_indices[BVH_N + 1];
  obox sub_boxes[BVH_N];
  multi_split(axes_minmax, boxes, indices, nboxes, sub_indices, sub_boxes, naxes);
  for (uint32_t i = 0; i < BVH_N; ++i) {
    const uint32_t sn = (uint32_t)(sub_indices[i + 1] - sub_indices[i]);
    if (sn == 1) nodes->a[nodei].child[i] = sub_indices[i][0];
  }
  /* Children are appended in slot order, each immediately before its own
   * subtree (the reference's parallel branch produces the same layout). */
  for (uint32_t i = 0; i < BVH_N; ++i) {
    const uint32_t sn = (uint32_t)(sub_indices[i + 1] - sub_indices[i]);
    if (sn != 1) {
      const uint32_t start = nv_push(nodes);
      nodes->a[nodei].child[i] = start | NODE_INTERNAL;
      init_node(nodes, start, &sub_boxes[i], boxes, sub_indices[i], sn, naxes);
    }
  }
}

/* BVH::init (4579-4634) over n boxes of 6 floats (x0 x1 y0 y1 z0 z1; a 2D box
 * has z0 = z1 = 0 and naxes = 2).  Returns the node count and the nodes
 * (4 uint32 each, malloc'ed, NULL for an empty tree). */
uint32_t oracle_bvh4_build(const float* boxes6, uint32_t nboxes, int naxes, uint32_t** nodes_out) {
  const obox* boxes = (const obox*)boxes6;
  *nodes_out = NULL;
  if (nboxes == 0) return 0;
  uint32_t n = nboxes;
  uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * n);
  for (uint32_t i = 0; i < n; ++i) idx[i] = i;
  /* utExcludeNaNInfBoxIndices (4535-4577) */
  uint32_t* end = idx + n;
  uint32_t* p = idx;
  for (; p != end; ++p)
    if (box_exclude(&boxes[*p])) break;
  if (p != end) {
    uint32_t* nan_start = p;
    for (++p; p != end; ++p)
      if (!box_exclude(&boxes[*p])) *nan_start++ = *p;
    n = (uint32_t)(nan_start - idx);
  }
  uint32_t count = 0;
  if (n > 0) {
    /* computeFullBoundingBox (4856-4913) over the kept boxes.  From 8192
     * boxes on a host with more than one hardware thread (the reference asks
     * std::thread::hardware_concurrency, 3670-3672) its parallel branch
     * starts each thread's box from UT_Array::setSize's zero fill of the POD
     * Box (2342-2360), so the origin is folded into the root bound. */
    obox full;
    if (n >= 2 * 4096 && get_nprocs() > 1) {
      for (int a = 0; a < 3; ++a) full.v[a][0] = full.v[a][1] = 0.0f;
      for (uint32_t i = 0; i < n; ++i) box_combine(&full, &boxes[idx[i]]);
    } else {
      full = boxes[idx[0]];
      for (uint32_t i = 1; i < n; ++i) box_combine(&full, &boxes[idx[i]]);
    }
    nodevec nvv = {NULL, 0, 0};
    nv_push(&nvv);
    init_node(&nvv, 0, &full, boxes, idx, n, naxes);
    *nodes_out = (uint32_t*)nvv.a;
    count = nvv.n;
  }
  free(idx);
  return count;
}

/* ---------------- precompute (6598-7128) ---------------- */
typedef struct {
  float maxP2[4];
  float P[3][4];
  float N[3][4];
  float NijDiag[3][4];
  float Nxy_Nyx[4], Nyz_Nzy[4], Nzx_Nxz[4];
  float NijkDiag[3][4];
  float SumPermuteNxyz[4];
  float N2xxy_yxx[4], N2xxz_zxx[4], N2yyz_zyy[4], N2yyx_xyy[4], N2zzx_xzz[4], N2zzy_yzz[4];
} boxdata_t; /* 92 floats, same order as BoxData (6477-6521) */

typedef struct {
  obox box;
  float averageP[3], areaP[3], N[3];
  float area;
  float NijDiag[3];
  float Nxy, Nyx, Nyz, Nzy, Nzx, Nxz;
  float NijkDiag[3];
  float SumPermuteNxyz, N2xxy_yxx, N2xxz_zxx, N2yyz_zyy, N2yyx_xyy, N2zzx_xzz, N2zzy_yzz;
} localdata_t; /* LocalData (6612-6640) */

struct oracle_soup {
  int32_t nv, nf, order;
  float* U;    /* nv x 3 */
  int32_t* F;  /* nf x 3 */
  obox* tri_boxes;
  onode* nodes;
  uint32_t nnodes;
  boxdata_t* data;
};

/* compute_integrals lambda (6760-6844): integral over the triangle (split at
 * the middle vertex along axis i) of (p_i-P_i)^2 and (p_i-P_i)(p_j-P_j). */
static void compute_integrals(const float* a, const float* b, const float* c, const float* P, float* integral_ii,
                              float* integral_ij, float* integral_ik, const int i) {
  const float oab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const float oac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const float ocb[3] = {b[0] - c[0], b[1] - c[1], b[2] - c[2]};
  const float t = oab[i] / oac[i];
  const int j = (i == 2) ? 0 : (i + 1);
  const int k = (j == 2) ? 0 : (j + 1);
  const float jdiff = t * oac[j] - oab[j];
  const float kdiff = t * oac[k] - oab[k];
  float cross_a[3], cross_c[3];
  cross_a[0] = (jdiff * oab[k] - kdiff * oab[j]);
  cross_a[1] = kdiff * oab[i];
  cross_a[2] = jdiff * oab[i];
  cross_c[0] = (jdiff * ocb[k] - kdiff * ocb[j]);
  cross_c[1] = kdiff * ocb[i];
  cross_c[2] = jdiff * ocb[i];
  const float area_scale_a = sqrtf(cross_a[0] * cross_a[0] + cross_a[1] * cross_a[1] + cross_a[2] * cross_a[2]);
  const float area_scale_c = sqrtf(cross_c[0] * cross_c[0] + cross_c[1] * cross_c[1] + cross_c[2] * cross_c[2]);
  const float Pai = a[i] - P[i];
  const float Pci = c[i] - P[i];
  const float int_ii_a =
      area_scale_a * (0.5f * Pai * Pai + (float)(2.0 / 3.0) * Pai * oab[i] + 0.25f * oab[i] * oab[i]);
  const float int_ii_c =
      area_scale_c * (0.5f * Pci * Pci + (float)(2.0 / 3.0) * Pci * ocb[i] + 0.25f * ocb[i] * ocb[i]);
  *integral_ii = int_ii_a + int_ii_c;
  int jk = j;
  float* integral = integral_ij;
  float diff = jdiff;
  for (;;) {
    if (integral) {
      const float obmidj = b[jk] + 0.5f * diff;
      const float oabmidj = obmidj - a[jk];
      const float ocbmidj = obmidj - c[jk];
      const float Paj = a[jk] - P[jk];
      const float Pcj = c[jk] - P[jk];
      const float int_ij_a = area_scale_a * (0.5f * Pai * Paj + (float)(1.0 / 3.0) * Pai * oabmidj +
                                             (float)(1.0 / 3.0) * Paj * oab[i] + 0.25f * oab[i] * oabmidj);
      const float int_ij_c = area_scale_c * (0.5f * Pci * Pcj + (float)(1.0 / 3.0) * Pci * ocbmidj +
                                             (float)(1.0 / 3.0) * Pcj * ocb[i] + 0.25f * ocb[i] * ocbmidj);
      *integral = int_ij_a + int_ij_c;
    }
    if (jk == k) break;
    jk = k;
    integral = integral_ik;
    diff = kdiff;
  }
}

static inline void sort3_axis(const float* vals[3], int axis, int order[3]) {
  order[0] = 0;
  order[1] = 1;
  order[2] = 2;
  int t;
  if (vals[0][axis] > vals[1][axis]) { t = order[0]; order[0] = order[1]; order[1] = t; }
  if (vals[order[0]][axis] > vals[2][axis]) { t = order[0]; order[0] = order[2]; order[2] = t; }
  if (vals[order[1]][axis] > vals[order[2]][axis]) { t = order[1]; order[1] = order[2]; order[2] = t; }
}

/* PrecomputeFunctors::item (6668-6890) */
static void pc_item(const oracle_soup* s, uint32_t itemi, localdata_t* d) {
  const int32_t* tp = s->F + 3 * (int64_t)itemi;
  const float* a = s->U + 3 * (int64_t)tp[0];
  const float* b = s->U + 3 * (int64_t)tp[1];
  const float* c = s->U + 3 * (int64_t)tp[2];
  const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  d->box = s->tri_boxes[itemi];
  const float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  float N[3] = {0.5f * cr[0], 0.5f * cr[1], 0.5f * cr[2]};
  const float area2 = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
  const float area = sqrtf(area2);
  const float third = 1.0f / 3.0f;
  float P[3];
  for (int i = 0; i < 3; ++i) P[i] = ((a[i] + b[i]) + c[i]) * third;
  for (int i = 0; i < 3; ++i) {
    d->averageP[i] = P[i];
    d->areaP[i] = P[i] * area;
    d->N[i] = N[i];
  }
  d->area = area;
  if (s->order < 1) return;
  d->NijDiag[0] = d->NijDiag[1] = d->NijDiag[2] = 0;
  d->Nxy = d->Nyx = d->Nyz = d->Nzy = d->Nzx = d->Nxz = 0;
  if (s->order < 2) return;
  if (area == 0) {
    d->NijkDiag[0] = d->NijkDiag[1] = d->NijkDiag[2] = 0;
    d->SumPermuteNxyz = d->N2xxy_yxx = d->N2xxz_zxx = d->N2yyz_zyy = 0;
    d->N2yyx_xyy = d->N2zzx_xzz = d->N2zzy_yzz = 0;
    return;
  }
  const float rarea = 1 / area;
  const float n[3] = {N[0] * rarea, N[1] * rarea, N[2] * rarea};
  const float* values[3] = {a, b, c};
  int ox[3], oy[3], oz[3];
  sort3_axis(values, 0, ox);
  const float dx = values[ox[2]][0] - values[ox[0]][0];
  sort3_axis(values, 1, oy);
  const float dy = values[oy[2]][1] - values[oy[0]][1];
  sort3_axis(values, 2, oz);
  const float dz = values[oz[2]][2] - values[oz[0]][2];
  float ixx = 0, ixy = 0, iyy = 0, iyz = 0, izz = 0, izx = 0;
  if (dx > 0)
    compute_integrals(values[ox[0]], values[ox[1]], values[ox[2]], P, &ixx, ((dx >= dy && dy > 0) ? &ixy : NULL),
                      ((dx >= dz && dz > 0) ? &izx : NULL), 0);
  if (dy > 0)
    compute_integrals(values[oy[0]], values[oy[1]], values[oy[2]], P, &iyy, ((dy >= dz && dz > 0) ? &iyz : NULL),
                      ((dx < dy && dx > 0) ? &ixy : NULL), 1);
  if (dz > 0)
    compute_integrals(values[oz[0]], values[oz[1]], values[oz[2]], P, &izz, ((dx < dz && dx > 0) ? &izx : NULL),
                      ((dy < dz && dy > 0) ? &iyz : NULL), 2);
  d->NijkDiag[0] = ixx * n[0];
  d->NijkDiag[1] = iyy * n[1];
  d->NijkDiag[2] = izz * n[2];
  d->SumPermuteNxyz = 2 * (n[0] * iyz + n[1] * izx + n[2] * ixy);
  const float Nxxy = n[0] * ixy, Nxxz = n[0] * izx, Nyyz = n[1] * iyz;
  const float Nyyx = n[1] * ixy, Nzzx = n[2] * izx, Nzzy = n[2] * iyz;
  d->N2xxy_yxx = 2 * Nxxy + n[1] * ixx;
  d->N2xxz_zxx = 2 * Nxxz + n[2] * ixx;
  d->N2yyz_zyy = 2 * Nyyz + n[2] * iyy;
  d->N2yyx_xyy = 2 * Nyyx + n[0] * iyy;
  d->N2zzx_xzz = 2 * Nzzx + n[0] * izz;
  d->N2zzy_yzz = 2 * Nzzy + n[1] * izz;
}

/* PrecomputeFunctors::post (6892-7118) */
static void pc_post(const oracle_soup* s, uint32_t nodei, localdata_t* out, int nchildren, const localdata_t* ch) {
  boxdata_t* bd = &s->data[nodei];
  float N[3] = {ch[0].N[0], ch[0].N[1], ch[0].N[2]};
  for (int k = 0; k < 3; ++k) {
    bd->N[k][0] = N[k];
    bd->P[k][0] = ch[0].averageP[k];
  }
  float areaP[3] = {ch[0].areaP[0], ch[0].areaP[1], ch[0].areaP[2]};
  float area = ch[0].area;
  for (int i = 1; i < nchildren; ++i) {
    for (int k = 0; k < 3; ++k) {
      N[k] += ch[i].N[k];
      bd->N[k][i] = ch[i].N[k];
      areaP[k] += ch[i].areaP[k];
      bd->P[k][i] = ch[i].averageP[k];
    }
    area += ch[i].area;
  }
  for (int i = nchildren; i < BVH_N; ++i)
    for (int k = 0; k < 3; ++k) bd->N[k][i] = bd->P[k][i] = 0;
  for (int k = 0; k < 3; ++k) {
    out->N[k] = N[k];
    out->areaP[k] = areaP[k];
  }
  out->area = area;
  obox box = ch[0].box;
  for (int i = 1; i < nchildren; ++i) box_combine(&box, &ch[i].box);
  float averageP[3];
  if (area > 0) {
    const float r = 1 / area;
    for (int k = 0; k < 3; ++k) averageP[k] = areaP[k] * r;
  } else {
    for (int k = 0; k < 3; ++k) averageP[k] = 0.5f * (box.v[k][0] + box.v[k][1]);
  }
  for (int k = 0; k < 3; ++k) out->averageP[k] = averageP[k];
  out->box = box;
  for (int i = 0; i < nchildren; ++i) {
    const obox* lb = &ch[i].box;
    const float* lp = ch[i].averageP;
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = hmax(lp[k] - lb->v[k][0], lb->v[k][1] - lp[k]);
    bd->maxP2[i] = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  }
  for (int i = nchildren; i < BVH_N; ++i) bd->maxP2[i] = INFINITY;
  if (s->order < 1) return;

  for (int k = 0; k < 3; ++k) out->NijDiag[k] = ch[0].NijDiag[k];
  out->Nxy = out->Nyx = out->Nyz = out->Nzy = out->Nzx = out->Nxz = 0;
  for (int k = 0; k < 3; ++k) out->NijkDiag[k] = ch[0].NijkDiag[k];
  out->SumPermuteNxyz = ch[0].SumPermuteNxyz;
  out->N2xxy_yxx = ch[0].N2xxy_yxx;
  out->N2xxz_zxx = ch[0].N2xxz_zxx;
  out->N2yyz_zyy = ch[0].N2yyz_zyy;
  out->N2yyx_xyy = ch[0].N2yyx_xyy;
  out->N2zzx_xzz = ch[0].N2zzx_xzz;
  out->N2zzy_yzz = ch[0].N2zzy_yzz;
  for (int i = 1; i < nchildren; ++i) {
    for (int k = 0; k < 3; ++k) out->NijDiag[k] += ch[i].NijDiag[k];
    for (int k = 0; k < 3; ++k) out->NijkDiag[k] += ch[i].NijkDiag[k];
    out->SumPermuteNxyz += ch[i].SumPermuteNxyz;
    out->N2xxy_yxx += ch[i].N2xxy_yxx;
    out->N2xxz_zxx += ch[i].N2xxz_zxx;
    out->N2yyz_zyy += ch[i].N2yyz_zyy;
    out->N2yyx_xyy += ch[i].N2yyx_xyy;
    out->N2zzx_xzz += ch[i].N2zzx_xzz;
    out->N2zzy_yzz += ch[i].N2zzy_yzz;
  }
  for (int i = 0; i < nchildren; ++i) {
    for (int k = 0; k < 3; ++k) bd->NijDiag[k][i] = ch[i].NijDiag[k];
    bd->Nxy_Nyx[i] = ch[i].Nxy + ch[i].Nyx;
    bd->Nyz_Nzy[i] = ch[i].Nyz + ch[i].Nzy;
    bd->Nzx_Nxz[i] = ch[i].Nzx + ch[i].Nxz;
    for (int k = 0; k < 3; ++k) bd->NijkDiag[k][i] = ch[i].NijkDiag[k];
    bd->SumPermuteNxyz[i] = ch[i].SumPermuteNxyz;
    bd->N2xxy_yxx[i] = ch[i].N2xxy_yxx;
    bd->N2xxz_zxx[i] = ch[i].N2xxz_zxx;
    bd->N2yyz_zyy[i] = ch[i].N2yyz_zyy;
    bd->N2yyx_xyy[i] = ch[i].N2yyx_xyy;
    bd->N2zzx_xzz[i] = ch[i].N2zzx_xzz;
    bd->N2zzy_yzz[i] = ch[i].N2zzy_yzz;
  }
  for (int i = nchildren; i < BVH_N; ++i) {
    for (int k = 0; k < 3; ++k) bd->NijDiag[k][i] = 0;
    bd->Nxy_Nyx[i] = bd->Nyz_Nzy[i] = bd->Nzx_Nxz[i] = 0;
    for (int k = 0; k < 3; ++k) bd->NijkDiag[k][i] = 0;
    bd->SumPermuteNxyz[i] = 0;
    bd->N2xxy_yxx[i] = bd->N2xxz_zxx[i] = bd->N2yyz_zyy[i] = 0;
    bd->N2yyx_xyy[i] = bd->N2zzx_xzz[i] = bd->N2zzy_yzz[i] = 0;
  }
  for (int i = 0; i < nchildren; ++i) {
    const localdata_t* c = &ch[i];
    float dp[3];
    for (int k = 0; k < 3; ++k) dp[k] = c->averageP[k] - out->averageP[k];
    const float* n = c->N;
    for (int k = 0; k < 3; ++k) out->NijDiag[k] += n[k] * dp[k];
    const float Nxy = c->Nxy + n[0] * dp[1];
    const float Nyx = c->Nyx + n[1] * dp[0];
    const float Nyz = c->Nyz + n[1] * dp[2];
    const float Nzy = c->Nzy + n[2] * dp[1];
    const float Nzx = c->Nzx + n[2] * dp[0];
    const float Nxz = c->Nxz + n[0] * dp[2];
    out->Nxy += Nxy;
    out->Nyx += Nyx;
    out->Nyz += Nyz;
    out->Nzy += Nzy;
    out->Nzx += Nzx;
    out->Nxz += Nxz;
    if (s->order >= 2) {
      for (int k = 0; k < 3; ++k) out->NijkDiag[k] += (2.0f * dp[k]) * c->NijDiag[k] + (dp[k] * dp[k]) * n[k];
      out->SumPermuteNxyz += (dp[0] * (Nyz + Nzy) + dp[1] * (Nzx + Nxz) + dp[2] * (Nxy + Nyx));
      out->N2xxy_yxx += 2 * (dp[1] * c->NijDiag[0] + dp[0] * c->Nxy + n[0] * dp[0] * dp[1]) + 2 * c->Nyx * dp[0] +
                        n[1] * dp[0] * dp[0];
      out->N2xxz_zxx += 2 * (dp[2] * c->NijDiag[0] + dp[0] * c->Nxz + n[0] * dp[0] * dp[2]) + 2 * c->Nzx * dp[0] +
                        n[2] * dp[0] * dp[0];
      out->N2yyz_zyy += 2 * (dp[2] * c->NijDiag[1] + dp[1] * c->Nyz + n[1] * dp[1] * dp[2]) + 2 * c->Nzy * dp[1] +
                        n[2] * dp[1] * dp[1];
      out->N2yyx_xyy += 2 * (dp[0] * c->NijDiag[1] + dp[1] * c->Nyx + n[1] * dp[1] * dp[0]) + 2 * c->Nxy * dp[1] +
                        n[0] * dp[1] * dp[1];
      out->N2zzx_xzz += 2 * (dp[0] * c->NijDiag[2] + dp[2] * c->Nzx + n[2] * dp[2] * dp[0]) + 2 * c->Nxz * dp[2] +
                        n[0] * dp[2] * dp[2];
      out->N2zzy_yzz += 2 * (dp[1] * c->NijDiag[2] + dp[2] * c->Nzy + n[2] * dp[2] * dp[1]) + 2 * c->Nyz * dp[2] +
                        n[1] * dp[2] * dp[2];
    }
  }
}

/* BVH::traverseHelper (4641-4667) with the precompute functors */
static void pc_traverse(const oracle_soup* s, uint32_t nodei, localdata_t* out) {
  localdata_t local[BVH_N];
  memset(local, 0, sizeof(local));
  const onode* node = &s->nodes[nodei];
  int sidx;
  for (sidx = 0; sidx < BVH_N; ++sidx) {
    const uint32_t ni = node->child[sidx];
    if (ni & NODE_INTERNAL) {
      if (ni == NODE_EMPTY) break;
      pc_traverse(s, ni & ~NODE_INTERNAL, &local[sidx]);
    } else {
      pc_item(s, ni, &local[sidx]);
    }
  }
  pc_post(s, nodei, out, sidx, local);
}

static oracle_soup* soup_build(const float* V, int32_t nv, const int32_t* F, int32_t nf, int order,
                               const uint32_t* nodes, int32_t nnodes);

oracle_soup* oracle_soup_build(const float* V, int32_t nv, const int32_t* F, int32_t nf, int order) {
  return soup_build(V, nv, F, nf, order, NULL, 0);
}

/* The moments (UT_SolidAngle::init's precompute, 6644-7118) over a given
 * 4-wide tree in the reference's node encoding, e.g. a device-built LBVH. */
oracle_soup* oracle_soup_build_nodes(const float* V, int32_t nv, const int32_t* F, int32_t nf, int order,
                                     const uint32_t* nodes, int32_t nnodes) {
  return soup_build(V, nv, F, nf, order, nodes, nnodes);
}

static oracle_soup* soup_build(const float* V, int32_t nv, const int32_t* F, int32_t nf, int order,
                               const uint32_t* nodes, int32_t nnodes) {
  oracle_soup* s = (oracle_soup*)calloc(1, sizeof(oracle_soup));
  s->nv = nv;
  s->nf = nf;
  s->order = order;
  s->U = (float*)malloc(sizeof(float) * 3 * (size_t)(nv > 0 ? nv : 1));
  s->F = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(nf > 0 ? nf : 1));
  memcpy(s->U, V, sizeof(float) * 3 * (size_t)nv);
  memcpy(s->F, F, sizeof(int32_t) * 3 * (size_t)nf);
  /* triangle boxes (6566-6590): initBounds(p0), enlargeBounds(p1), enlargeBounds(p2) */
  s->tri_boxes = (obox*)malloc(sizeof(obox) * (size_t)(nf > 0 ? nf : 1));
  for (int32_t i = 0; i < nf; ++i) {
    obox* b = &s->tri_boxes[i];
    const float* p0 = s->U + 3 * (int64_t)s->F[3 * i];
    for (int a = 0; a < 3; ++a) b->v[a][0] = b->v[a][1] = p0[a];
    for (int k = 1; k < 3; ++k) {
      const float* p = s->U + 3 * (int64_t)s->F[3 * i + k];
      for (int a = 0; a < 3; ++a) {
        b->v[a][0] = hmin(b->v[a][0], p[a]);
        b->v[a][1] = hmax(b->v[a][1], p[a]);
      }
    }
  }
  if (nodes) {
    s->nnodes = (uint32_t)nnodes;
    s->nodes = (onode*)malloc(sizeof(onode) * (size_t)(nnodes > 0 ? nnodes : 1));
    memcpy(s->nodes, nodes, sizeof(onode) * (size_t)nnodes);
  } else {
    s->nnodes = oracle_bvh4_build((const float*)s->tri_boxes, (uint32_t)nf, 3, (uint32_t**)&s->nodes);
  }
  s->data = (boxdata_t*)calloc(s->nnodes > 0 ? s->nnodes : 1, sizeof(boxdata_t));
  if (s->nnodes > 0) {
    localdata_t root;
    memset(&root, 0, sizeof(root));
    pc_traverse(s, 0, &root);
  }
  return s;
}

void oracle_soup_free(oracle_soup* s) {
  if (!s) return;
  free(s->U);
  free(s->F);
  free(s->tri_boxes);
  free(s->nodes);
  free(s->data);
  free(s);
}

int32_t oracle_soup_num_nodes(const oracle_soup* s) { return (int32_t)s->nnodes; }
void oracle_soup_nodes(const oracle_soup* s, uint32_t* out) {
  if (s->nnodes) memcpy(out, s->nodes, sizeof(onode) * s->nnodes);
}
void oracle_soup_boxdata(const oracle_soup* s, float* out) {
  if (s->nnodes) memcpy(out, s->data, sizeof(boxdata_t) * s->nnodes);
}

/* UTsignedSolidAngleTri (6070-6110) in float */
static float solid_angle_tri(const float* a, const float* b, const float* c, const float* q) {
  float qa[3], qb[3], qc[3];
  for (int i = 0; i < 3; ++i) {
    qa[i] = a[i] - q[i];
    qb[i] = b[i] - q[i];
    qc[i] = c[i] - q[i];
  }
  const float al = sqrtf(qa[0] * qa[0] + qa[1] * qa[1] + qa[2] * qa[2]);
  const float bl = sqrtf(qb[0] * qb[0] + qb[1] * qb[1] + qb[2] * qb[2]);
  const float cl = sqrtf(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2]);
  if (al == 0 || bl == 0 || cl == 0) return 0.0f;
  const float ra = 1 / al, rb = 1 / bl, rc = 1 / cl;
  for (int i = 0; i < 3; ++i) {
    qa[i] *= ra;
    qb[i] *= rb;
    qc[i] *= rc;
  }
  const float u[3] = {qb[0] - qa[0], qb[1] - qa[1], qb[2] - qa[2]};
  const float w[3] = {qc[0] - qa[0], qc[1] - qa[1], qc[2] - qa[2]};
  const float cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
  const float num = qa[0] * cr[0] + qa[1] * cr[1] + qa[2] * cr[2];
  if (num == 0) return 0.0f;
  const float den = 1.0f + (qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2]) +
                    (qa[0] * qc[0] + qa[1] * qc[1] + qa[2] * qc[2]) + (qb[0] * qc[0] + qb[1] * qc[1] + qb[2] * qc[2]);
  return 2.0f * atan2f(num, den);
}

typedef struct {
  const oracle_soup* s;
  float q[3];
  float acc2;
  int64_t visits, far_accepted, exact;
  float near_d;  /* < 0: near triangles not counted */
  int64_t near;  /* exact triangles with |q - P| < sqrt(maxP2) + near_d */
} sa_ctx;

/* SolidAngleFunctors::pre (7176-7258): far-field Taylor terms of the child
 * slots not descended into; returns the descend bitmask. */
static uint32_t sa_pre(sa_ctx* c, uint32_t nodei, float* data_for_parent) {
  const boxdata_t* d = &c->s->data[nodei];
  c->visits++;
  float q[3][4], ql2[4];
  uint32_t descend = 0;
  for (int l = 0; l < 4; ++l) {
    for (int k = 0; k < 3; ++k) q[k][l] = c->q[k] - d->P[k][l];
    ql2[l] = q[0][l] * q[0][l] + q[1][l] * q[1][l] + q[2][l] * q[2][l];
    if (ql2[l] <= d->maxP2[l] * c->acc2) descend |= 1u << l;
  }
  if (descend == 0xFu) {
    *data_for_parent = 0;
    return 0xFu;
  }
  float om[4];
  for (int l = 0; l < 4; ++l) {
    const float m2 = 1.0f / ql2[l];
    const float m1 = sqrtf(m2);
    float x = q[0][l] * m1, y = q[1][l] * m1, z = q[2][l] * m1;
    float o = -m2 * (x * d->N[0][l] + y * d->N[1][l] + z * d->N[2][l]);
    if (c->s->order >= 1) {
      const float x2 = x * x, y2 = y * y, z2 = z * z;
      const float m3 = m2 * m1;
      const float o1 = m3 * (d->NijDiag[0][l] + d->NijDiag[1][l] + d->NijDiag[2][l] -
                             3.0f * ((x2 * d->NijDiag[0][l] + y2 * d->NijDiag[1][l] + z2 * d->NijDiag[2][l]) +
                                     x * y * d->Nxy_Nyx[l] + x * z * d->Nzx_Nxz[l] + y * z * d->Nyz_Nzy[l]));
      o += o1;
      if (c->s->order >= 2) {
        const float x3 = x2 * x, y3 = y2 * y, z3 = z2 * z;
        const float m4 = m2 * m2;
        const float t0[3] = {d->N2yyx_xyy[l] + d->N2zzx_xzz[l], d->N2zzy_yzz[l] + d->N2xxy_yxx[l],
                             d->N2xxz_zxx[l] + d->N2yyz_zyy[l]};
        const float t1[3] = {y * d->N2xxy_yxx[l] + z * d->N2xxz_zxx[l], z * d->N2yyz_zyy[l] + x * d->N2yyx_xyy[l],
                             x * d->N2zzx_xzz[l] + y * d->N2zzy_yzz[l]};
        const float s3[3] = {3.0f * d->NijkDiag[0][l] + t0[0], 3.0f * d->NijkDiag[1][l] + t0[1],
                             3.0f * d->NijkDiag[2][l] + t0[2]};
        const float o2 =
            m4 * (1.5f * (x * s3[0] + y * s3[1] + z * s3[2]) -
                  7.5f * ((x3 * d->NijkDiag[0][l] + y3 * d->NijkDiag[1][l] + z3 * d->NijkDiag[2][l]) +
                          x * y * z * d->SumPermuteNxyz[l] + (x2 * t1[0] + y2 * t1[1] + z2 * t1[2])));
        o += o2;
      }
    }
    om[l] = o;
  }
  uint32_t mask = 0;
  for (int l = 0; l < 4; ++l) {
    int finite = isfinite(om[l]) ? 1 : 0;
    if (finite && !((descend >> l) & 1u)) mask |= 1u << l;
    if (!((mask >> l) & 1u)) om[l] = 0.0f; /* bitwise AND with a zero mask gives +0 */
  }
  descend = (~mask) & 0xFu;
  for (int l = 0; l < 4; ++l)
    if ((mask >> l) & 1u) c->far_accepted++;
  float sum = om[0];
  for (int l = 1; l < 4; ++l) sum += om[l];
  *data_for_parent = sum;
  return descend;
}

/* BVH::traverseVectorHelper (4818-4848) + SolidAngleFunctors item/post (7260-7278) */
static void sa_traverse(sa_ctx* c, uint32_t nodei, float* data_for_parent) {
  const onode* node = &c->s->nodes[nodei];
  uint32_t descend = sa_pre(c, nodei, data_for_parent);
  if (!descend) return;
  float local[BVH_N];
  int sidx;
  for (sidx = 0; sidx < BVH_N; ++sidx) {
    if ((descend >> sidx) & 1u) {
      const uint32_t ni = node->child[sidx];
      if (ni & NODE_INTERNAL) {
        if (ni == NODE_EMPTY) {
          descend &= (1u << sidx) - 1;
          break;
        }
        sa_traverse(c, ni & ~NODE_INTERNAL, &local[sidx]);
      } else {
        const int32_t* tp = c->s->F + 3 * (int64_t)ni;
        c->exact++;
        if (c->near_d >= 0.0f) {
          const boxdata_t* d = &c->s->data[nodei];
          const float x = c->q[0] - d->P[0][sidx], y = c->q[1] - d->P[1][sidx], z = c->q[2] - d->P[2][sidx];
          const float rr = sqrtf(d->maxP2[sidx]) + c->near_d;
          if (x * x + y * y + z * z < rr * rr) c->near++;
        }
        local[sidx] = solid_angle_tri(c->s->U + 3 * (int64_t)tp[0], c->s->U + 3 * (int64_t)tp[1],
                                      c->s->U + 3 * (int64_t)tp[2], c->q);
      }
    }
  }
  float sum = (descend & 1u) ? local[0] : 0.0f;
  for (int i = 1; i < sidx; ++i) sum += ((descend >> i) & 1u) ? local[i] : 0.0f;
  *data_for_parent += sum;
}

float oracle_soup_solid_angle(const oracle_soup* s, const float* q, float accuracy_scale) {
  /* empty tree: the reference returns an uninitialised value (7281-7283,
   * 4810-4811); defined here as 0 */
  if (s->nnodes == 0) return 0.0f;
  sa_ctx c;
  c.s = s;
  c.q[0] = q[0];
  c.q[1] = q[1];
  c.q[2] = q[2];
  c.acc2 = accuracy_scale * accuracy_scale;
  c.visits = c.far_accepted = c.exact = c.near = 0;
  c.near_d = -1.0f;
  float sum;
  sa_traverse(&c, 0, &sum);
  return sum;
}

void oracle_soup_count_near(const oracle_soup* s, const float* q, float accuracy_scale, float near_d,
                            int64_t* visits, int64_t* far_accepted, int64_t* exact_tris, int64_t* near_tris) {
  *visits = *far_accepted = *exact_tris = *near_tris = 0;
  if (s->nnodes == 0) return;
  sa_ctx c;
  c.s = s;
  c.q[0] = q[0];
  c.q[1] = q[1];
  c.q[2] = q[2];
  c.acc2 = accuracy_scale * accuracy_scale;
  c.visits = c.far_accepted = c.exact = c.near = 0;
  c.near_d = near_d;
  float sum;
  sa_traverse(&c, 0, &sum);
  *visits = c.visits;
  *far_accepted = c.far_accepted;
  *exact_tris = c.exact;
  *near_tris = c.near;
}

void oracle_soup_count(const oracle_soup* s, const float* q, float accuracy_scale, int64_t* visits,
                       int64_t* far_accepted, int64_t* exact_tris) {
  int64_t near;
  oracle_soup_count_near(s, q, accuracy_scale, -1.0f, visits, far_accepted, exact_tris, &near);
}

typedef struct {
  const oracle_soup* s;
  const double* Q;
  float acc;
  double* W;
} sq_ctx;

#define ORACLE_PI 3.1415926535897932384626433832795

static void sq_range(int64_t b, int64_t e, void* vctx) {
  sq_ctx* c = (sq_ctx*)vctx;
  for (int64_t p = b; p < e; ++p) {
    float q[3] = {(float)c->Q[3 * p], (float)c->Q[3 * p + 1], (float)c->Q[3 * p + 2]};
    c->W[p] = (double)oracle_soup_solid_angle(c->s, q, c->acc) / (4.0 * ORACLE_PI);
  }
}

/* fast_winding_number.cpp:418-437: parallel_for(nq, ..., 1000) */
void oracle_soup_query(const oracle_soup* s, const double* Q, int64_t nq, float accuracy_scale, double* W,
                       int nthreads) {
  sq_ctx c = {s, Q, accuracy_scale, W};
  oracle_parallel_for(nq, sq_range, &c, 1000, nthreads);
}
