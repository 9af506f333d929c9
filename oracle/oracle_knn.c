/*
 * oracle_knn.c -- CPU restatement of igl::knn (knn.cpp:24-127).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   oracle_knn      knn.cpp:72-127: per query one priority queue holding
 *                   points (entries 0 .. nv-1) and octree cells (nv + cell),
 *                   nearest first -- a point by |V(j) - p|, a cell by the
 *                   bound of knn.cpp:57-70 (see cube_dist) -- popped
 *                   until k points are out; a popped leaf pushes its first
 *                   point, a popped split cell its 8 children.  The queue is
 *                   a binary heap on (distance, entry), so equal distances pop
 *                   in ascending entry order (the reference leaves that order
 *                   to std::priority_queue).
 *                   knn.cpp:46-55 (nv <= k): every point of V by distance,
 *                   equal distances in ascending index order.
 *   oracle_knn_bf   brute force over a candidate list: the k nearest by
 *                   (squared distance, index).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
  double d;
  int64_t e;
} heap_item;

static int item_less(const heap_item* a, const heap_item* b) { return a->d < b->d || (a->d == b->d && a->e < b->e); }

typedef struct {
  heap_item* v;
  int64_t n, cap;
} heap;

static int heap_push(heap* h, double d, int64_t e) {
  if (h->n == h->cap) {
    const int64_t c = h->cap ? 2 * h->cap : 64;
    heap_item* nv = (heap_item*)realloc(h->v, sizeof(heap_item) * (size_t)c);
    if (!nv) return -1;
    h->v = nv;
    h->cap = c;
  }
  int64_t i = h->n++;
  h->v[i].d = d;
  h->v[i].e = e;
  while (i > 0) {
    const int64_t p = (i - 1) / 2;
    if (!item_less(&h->v[i], &h->v[p])) break;
    const heap_item t = h->v[i];
    h->v[i] = h->v[p];
    h->v[p] = t;
    i = p;
  }
  return 0;
}

static heap_item heap_pop(heap* h) {
  const heap_item top = h->v[0];
  h->v[0] = h->v[--h->n];
  int64_t i = 0;
  for (;;) {
    const int64_t l = 2 * i + 1, r = l + 1;
    int64_t m = i;
    if (l < h->n && item_less(&h->v[l], &h->v[m])) m = l;
    if (r < h->n && item_less(&h->v[r], &h->v[m])) m = r;
    if (m == i) break;
    const heap_item t = h->v[i];
    h->v[i] = h->v[m];
    h->v[m] = t;
    i = m;
  }
  return top;
}

/* |V(j) - p| summed as Eigen sums a row: (x^2 + y^2) + z^2 */
static double point_dist(const double* v, const double* p) {
  const double dx = v[0] - p[0], dy = v[1] - p[1], dz = v[2] - p[2];
  return sqrt((dx * dx + dy * dy) + dz * dz);
}

/* knn.cpp:57-70: W * |max(|(p - CN)/W| - 1, 0)|, evaluated as
 * |max(|p - CN| - W, 0)|, equal in exact arithmetic.  The reference's form
 * overflows for the tiny widths of the cells that split coincident points
 * (igl::octree splits them down to depth 30000): once W is below ~1e-154 of
 * |p - CN| its bound is infinite, and NaN at W = 0, so its search reaches such
 * a point only after every other one.  This form keeps a true lower bound, and
 * the search returns the k nearest candidates. */
static double cube_dist(const double* p, const double* c, double w) {
  double s = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double e = fabs(p[a] - c[a]) - w > 0.0 ? fabs(p[a] - c[a]) - w : 0.0;
    s += e * e;
  }
  return sqrt(s);
}

typedef struct {
  const double* P;
  const double* V;
  int64_t nv, k;
  const oracle_octree_t* t;
  int32_t* I;
  int failed;
} knn_ctx;

static void knn_range(int64_t b, int64_t e, void* vctx) {
  knn_ctx* x = (knn_ctx*)vctx;
  const double* CN = (const double*)x->t->CN;
  const double* W = (const double*)x->t->W;
  heap h = {NULL, 0, 0};
  for (int64_t i = b; i < e; ++i) {
    const double* p = x->P + 3 * i;
    h.n = 0;
    if (heap_push(&h, cube_dist(p, CN, W[0]), x->nv) != 0) {
      x->failed = 1;
      break;
    }
    int64_t found = 0;
    while (found < x->k) {
      if (h.n == 0) {  /* fewer reachable points than k (undefined in the reference) */
        x->failed = 1;
        break;
      }
      const heap_item top = heap_pop(&h);
      if (top.e < x->nv) {
        x->I[i * x->k + found++] = (int32_t)top.e;
      } else {
        const int64_t c = top.e - x->nv;
        const int32_t* ch = x->t->CH + 8 * c;
        if (ch[0] == -1) {
          if (x->t->cell_count[c] > 0) {
            const int64_t j = x->t->point_list[x->t->cell_start[c]];
            if (heap_push(&h, point_dist(x->V + 3 * j, p), j) != 0) x->failed = 1;
          }
        } else {
          for (int s = 0; s < 8; ++s)
            if (heap_push(&h, cube_dist(p, CN + 3 * (int64_t)ch[s], W[ch[s]]), x->nv + ch[s]) != 0) x->failed = 1;
        }
      }
    }
  }
  free(h.v);
}

typedef struct {
  const double* P;
  const double* V;
  int64_t nv, k;
  const int32_t* cand;
  int64_t ncand;
  int32_t* I;
  int failed;
} bf_ctx;

static void bf_range(int64_t b, int64_t e, void* vctx) {
  bf_ctx* x = (bf_ctx*)vctx;
  double* d = (double*)malloc(sizeof(double) * (size_t)(x->k + 1));
  int32_t* id = (int32_t*)malloc(sizeof(int32_t) * (size_t)(x->k + 1));
  if (!d || !id) {
    x->failed = 1;
    free(d);
    free(id);
    return;
  }
  for (int64_t i = b; i < e; ++i) {
    const double* p = x->P + 3 * i;
    int64_t m = 0; /* sorted ascending by (d2, index) */
    for (int64_t c = 0; c < x->ncand; ++c) {
      const int32_t j = x->cand ? x->cand[c] : (int32_t)c;
      const double* v = x->V + 3 * (int64_t)j;
      const double dx = v[0] - p[0], dy = v[1] - p[1], dz = v[2] - p[2];
      const double d2 = (dx * dx + dy * dy) + dz * dz;
      if (m == x->k && !(d2 < d[m - 1] || (d2 == d[m - 1] && j < id[m - 1]))) continue;
      int64_t s = m < x->k ? m++ : m - 1;
      while (s > 0 && (d2 < d[s - 1] || (d2 == d[s - 1] && j < id[s - 1]))) {
        d[s] = d[s - 1];
        id[s] = id[s - 1];
        --s;
      }
      d[s] = d2;
      id[s] = j;
    }
    for (int64_t s = 0; s < x->k; ++s) x->I[i * x->k + s] = s < m ? id[s] : -1;
  }
  free(d);
  free(id);
}

int oracle_knn_bf(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand,
                  int64_t ncand, int32_t* I, int nthreads) {
  if (k < 1) return -1;
  if (nv <= k) { /* knn.cpp:46-55 */
    cand = NULL;
    ncand = nv;
    k = nv;
  } else if (!cand) {
    ncand = nv;
  }
  if (ncand < k) return -1;
  bf_ctx x = {P, V, nv, k, cand, ncand, I, 0};
  oracle_parallel_for(np, bf_range, &x, 1000, nthreads);
  return x.failed ? -2 : 0;
}

int oracle_knn(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const oracle_octree_t* t,
               int32_t* I, int nthreads) {
  if (k < 1) return -1;
  if (nv <= k) return oracle_knn_bf(P, np, V, nv, k, NULL, 0, I, nthreads);
  knn_ctx x = {P, V, nv, k, t, I, 0};
  oracle_parallel_for(np, knn_range, &x, 1000, nthreads);
  return x.failed ? -1 : 0;
}
