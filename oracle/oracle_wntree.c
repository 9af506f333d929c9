/*
 * oracle_wntree.c -- TEST INFRASTRUCTURE (the checker, never the product).
 *
 * Restates the hierarchy igl::winding_number(V,F,O,W) builds for triangle
 * meshes (winding_number.cpp:39-53): igl::WindingNumberAABB over
 * igl::WindingNumberTree (WindingNumberAABB.h:149-310, WindingNumberTree.h:
 * 150-336), step by step and in the reference's order:
 *   - remove_duplicate_vertices(V,F,0) (remove_duplicate_vertices.cpp): the
 *     unique rows of V in lexicographic order (unique_rows.cpp, sortrows.cpp),
 *     F re-indexed into them;
 *   - every node's cap: exterior_edges (exterior_edges.cpp:46-100, edges of
 *     oriented_facets.cpp:36-43 sorted per row by sort2, sort.cpp:162-200,
 *     unique rows in lexicographic order, a signed count per edge, each
 *     non-zero edge repeated |count| times, oriented by the count's sign)
 *     closed by triangle_fan (triangle_fan.cpp:15-47) around the first vertex
 *     of the edge rand() % #edges picks, skipping edges that touch it;
 *   - grow(): no split at <= 100 faces or when the cap is not smaller than
 *     the faces (#cap - 2 >= #F); else split at the median (median.cpp,
 *     std::nth_element twice) of the face barycenters (barycenter.cpp) along
 *     the longest side of the node's box; faces with barycenter <= median go
 *     left; nodes are constructed (and call rand()) in depth-first pre-order;
 *   - winding_number(p): inside the box (closed interval) a node sums its two
 *     children (0 + left + right) or, as a leaf, its faces; outside it sums
 *     its cap while #cap - 2 < #F, else its faces; each face adds
 *     solid_angle(...) (solid_angle.cpp, already divided by 2 pi).
 * rand() is the C library's, called in the reference's order, so a build
 * consumes the same rand() values the reference's would from the same state.
 * std::nth_element is the C++ library's; its algorithm (GCC libstdc++ 11,
 * bits/stl_algo.h __introselect / __unguarded_partition_pivot /
 * __insertion_sort, bits/stl_heap.h for the heap fallback) is restated below
 * because the reference reads element n after a second nth_element call,
 * whose value depends on that algorithm.
 * Bitwise equality with the Eigen-compiled reference is not pinned (Eigen's
 * reduction order in .norm() is not restated; Eigen is absent here).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------- std::nth_element (libstdc++ introselect) ---------------- */
static void dswap(double* a, double* b) {
  double t = *a;
  *a = *b;
  *b = t;
}

static void move_median_to_first(double* result, double* a, double* b, double* c) {
  if (*a < *b) {
    if (*b < *c)
      dswap(result, b);
    else if (*a < *c)
      dswap(result, c);
    else
      dswap(result, a);
  } else if (*a < *c) {
    dswap(result, a);
  } else if (*b < *c) {
    dswap(result, c);
  } else {
    dswap(result, b);
  }
}

static double* unguarded_partition(double* first, double* last, double* pivot) {
  for (;;) {
    while (*first < *pivot) ++first;
    --last;
    while (*pivot < *last) --last;
    if (!(first < last)) return first;
    dswap(first, last);
    ++first;
  }
}

static double* unguarded_partition_pivot(double* first, double* last) {
  double* mid = first + (last - first) / 2;
  move_median_to_first(first, first + 1, mid, last - 1);
  return unguarded_partition(first + 1, last, first);
}

static void insertion_sort(double* first, double* last) {
  if (first == last) return;
  for (double* i = first + 1; i != last; ++i) {
    const double val = *i;
    if (val < *first) {
      memmove(first + 1, first, sizeof(double) * (size_t)(i - first));
      *first = val;
    } else {
      double* hole = i;
      double* next = i - 1;
      while (val < *next) {
        *hole = *next;
        hole = next;
        --next;
      }
      *hole = val;
    }
  }
}

/* bits/stl_heap.h: __adjust_heap with __push_heap, __make_heap, __pop_heap */
static void push_heap_(double* first, ptrdiff_t hole, ptrdiff_t top, double value) {
  ptrdiff_t parent = (hole - 1) / 2;
  while (hole > top && first[parent] < value) {
    first[hole] = first[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  first[hole] = value;
}

static void adjust_heap(double* first, ptrdiff_t hole, ptrdiff_t len, double value) {
  const ptrdiff_t top = hole;
  ptrdiff_t child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (first[child] < first[child - 1]) child--;
    first[hole] = first[child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    first[hole] = first[child - 1];
    hole = child - 1;
  }
  push_heap_(first, hole, top, value);
}

static void make_heap_(double* first, double* last) {
  const ptrdiff_t len = last - first;
  if (len < 2) return;
  ptrdiff_t parent = (len - 2) / 2;
  for (;;) {
    adjust_heap(first, parent, len, first[parent]);
    if (parent == 0) return;
    parent--;
  }
}

static void pop_heap_(double* first, double* last, double* result) {
  const double value = *result;
  *result = *first;
  adjust_heap(first, 0, last - first, value);
}

static void heap_select(double* first, double* middle, double* last) {
  make_heap_(first, middle);
  for (double* i = middle; i < last; ++i)
    if (*i < *first) pop_heap_(first, middle, i);
}

static int lg2(ptrdiff_t n) {
  int k = 0;
  while (n > 1) {
    n >>= 1;
    ++k;
  }
  return k;
}

static void nth_element_(double* first, double* nth, double* last) {
  if (first == last || nth == last) return;
  ptrdiff_t depth = 2 * lg2(last - first);
  while (last - first > 3) {
    if (depth == 0) {
      heap_select(first, nth + 1, last);
      dswap(first, nth);
      return;
    }
    --depth;
    double* cut = unguarded_partition_pivot(first, last);
    if (cut <= nth)
      first = cut;
    else
      last = cut;
  }
  insertion_sort(first, last);
}

/* median.cpp: n = size / 2; nth_element at n; for an even size a second
 * nth_element at n - 1 and m = 0.5 (v[n] + v[n-1]) */
static double median_(double* v, int64_t size) {
  const int64_t n = size / 2;
  nth_element_(v, v + n, v + size);
  if (size % 2 == 0) {
    nth_element_(v, v + n - 1, v + size);
    return 0.5 * (v[n] + v[n - 1]);
  }
  return v[n];
}

/* ---------------- exterior edges and the triangle fan ---------------- */
typedef struct {
  uint64_t key; /* (min << 32) | max of the edge's vertex indices */
  int sign;     /* +1 when the facet edge already had min first (sort2 swaps only on a > b) */
} ekey;

static int cmp_ekey(const void* x, const void* y) {
  const uint64_t a = ((const ekey*)x)->key, b = ((const ekey*)y)->key;
  return a < b ? -1 : (a > b ? 1 : 0);
}

/* exterior_edges.cpp:46-100 -> E (ne x 2), in unique-edge order */
static int32_t* exterior_edges(const int32_t* F, int64_t m, int64_t* ne) {
  ekey* k = (ekey*)malloc(sizeof(ekey) * (size_t)(3 * m + 1));
  for (int c = 0; c < 3; ++c)
    for (int64_t f = 0; f < m; ++f) {
      /* oriented_facets.cpp:36-43: rows (1,2), then (2,0), then (0,1) */
      const int32_t u = F[3 * f + (c + 1) % 3], v = F[3 * f + (c + 2) % 3];
      const int up = !(u > v);
      const uint32_t lo = (uint32_t)(up ? u : v), hi = (uint32_t)(up ? v : u);
      k[c * m + f].key = ((uint64_t)lo << 32) | hi;
      k[c * m + f].sign = up ? 1 : -1;
    }
  qsort(k, (size_t)(3 * m), sizeof(ekey), cmp_ekey);
  int32_t* E = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)(3 * m + 1));
  int64_t e = 0;
  for (int64_t a = 0; a < 3 * m;) {
    int64_t b = a;
    int64_t count = 0;
    while (b < 3 * m && k[b].key == k[a].key) count += k[b++].sign;
    const int32_t lo = (int32_t)(k[a].key >> 32), hi = (int32_t)(k[a].key & 0xFFFFFFFFu);
    const int64_t reps = count < 0 ? -count : count;
    for (int64_t r = 0; r < reps; ++r) {
      E[2 * e] = count > 0 ? lo : hi;
      E[2 * e + 1] = count > 0 ? hi : lo;
      ++e;
    }
    a = b;
  }
  free(k);
  *ne = e;
  return E;
}

/* triangle_fan.cpp:15-47 -> cap (ncap x 3) around the apex *s; no rand()
 * call (and *s = -1) without edges */
static int32_t* triangle_fan(const int32_t* E, int64_t ne, int64_t* ncap, int32_t* apex) {
  *ncap = 0;
  *apex = -1;
  if (ne == 0) return NULL;
  const int32_t s = E[2 * ((int64_t)rand() % ne)];
  *apex = s;
  int32_t* cap = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)ne);
  int64_t c = 0;
  for (int64_t i = 0; i < ne; ++i) {
    if (E[2 * i] == s || E[2 * i + 1] == s) continue;
    cap[3 * c] = s;
    cap[3 * c + 1] = E[2 * i];
    cap[3 * c + 2] = E[2 * i + 1];
    ++c;
  }
  *ncap = c;
  return cap;
}

/* ---------------- the tree ---------------- */
typedef struct {
  int64_t nf, ncap;
  int32_t* F;
  int32_t* cap;
  double lo[3], hi[3];
  int64_t child[2];
  int32_t apex; /* the fan's first vertex, -1 without edges */
} wnode;

struct oracle_wntree {
  double* SV;
  int64_t nsv;
  wnode* nodes;
  int64_t nnodes, capacity;
};

/* WindingNumberTree(parent, F) + WindingNumberAABB::init: cap, then the box
 * of the node's vertices (compute_min_max_corners, WindingNumberAABB.h:255-294) */
static int64_t node_new(oracle_wntree* t, int32_t* F, int64_t nf) {
  if (t->nnodes == t->capacity) {
    t->capacity = t->capacity ? 2 * t->capacity : 64;
    t->nodes = (wnode*)realloc(t->nodes, sizeof(wnode) * (size_t)t->capacity);
  }
  const int64_t id = t->nnodes++;
  wnode* n = &t->nodes[id];
  memset(n, 0, sizeof(*n));
  n->F = F;
  n->nf = nf;
  n->child[0] = n->child[1] = -1;
  int64_t ne = 0;
  int32_t* E = exterior_edges(F, nf, &ne);
  n->cap = triangle_fan(E, ne, &n->ncap, &n->apex);
  free(E);
  for (int d = 0; d < 3; ++d) {
    n->lo[d] = INFINITY;
    n->hi[d] = -INFINITY;
  }
  for (int64_t i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) {
        const double v = t->SV[3 * (int64_t)F[3 * i + j] + d];
        n->lo[d] = v < n->lo[d] ? v : n->lo[d];
        n->hi[d] = v > n->hi[d] ? v : n->hi[d];
      }
  return id;
}

/* WindingNumberAABB::grow, WindingNumberAABB.h:149-235 */
static void grow(oracle_wntree* t, int64_t id) {
  const int64_t nf = t->nodes[id].nf, ncap = t->nodes[id].ncap;
  if (nf <= 100 || (ncap - 2) >= nf) return;
  int max_d = -1;
  double max_len = -INFINITY;
  for (int d = 0; d < 3; ++d) {
    const double len = t->nodes[id].hi[d] - t->nodes[id].lo[d];
    if (len > max_len) {
      max_len = len;
      max_d = d;
    }
  }
  const int32_t* F = t->nodes[id].F;
  double* bc = (double*)malloc(sizeof(double) * (size_t)nf);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)nf);
  for (int64_t i = 0; i < nf; ++i) {
    double s = 0.0;  /* barycenter.cpp: zero, add the three corners, divide by 3 */
    for (int j = 0; j < 3; ++j) s += t->SV[3 * (int64_t)F[3 * i + j] + max_d];
    bc[i] = s / 3.0;
    tmp[i] = bc[i];
  }
  const double split = median_(tmp, nf);
  free(tmp);
  int64_t nl = 0;
  for (int64_t i = 0; i < nf; ++i) nl += bc[i] <= split;
  const int64_t nr = nf - nl;
  if (nl == 0 || nr == 0) {
    free(bc);
    return;
  }
  int32_t* LF = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)nl);
  int32_t* RF = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)nr);
  int64_t a = 0, b = 0;
  for (int64_t i = 0; i < nf; ++i) {
    if (bc[i] <= split)
      memcpy(LF + 3 * a++, F + 3 * i, 3 * sizeof(int32_t));
    else
      memcpy(RF + 3 * b++, F + 3 * i, 3 * sizeof(int32_t));
  }
  free(bc);
  const int64_t l = node_new(t, LF, nl);
  grow(t, l);
  const int64_t r = node_new(t, RF, nr);
  grow(t, r);
  t->nodes[id].child[0] = l;
  t->nodes[id].child[1] = r;
}

static const double* g_sort_v;
static int cmp_vrow(const void* x, const void* y) {
  const double* a = g_sort_v + 3 * (int64_t)(*(const int32_t*)x);
  const double* b = g_sort_v + 3 * (int64_t)(*(const int32_t*)y);
  for (int c = 0; c < 3; ++c) {
    if (a[c] < b[c]) return -1;
    if (b[c] < a[c]) return 1;
  }
  return 0;
}

oracle_wntree* oracle_wntree_build(const double* V, int64_t nv, const int32_t* F, int64_t nf) {
  oracle_wntree* t = (oracle_wntree*)calloc(1, sizeof(oracle_wntree));
  /* remove_duplicate_vertices(V,F,0): unique rows in sorted order */
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  for (int64_t i = 0; i < nv; ++i) idx[i] = (int32_t)i;
  g_sort_v = V;
  qsort(idx, (size_t)nv, sizeof(int32_t), cmp_vrow);
  int32_t* ic = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  t->SV = (double*)malloc(sizeof(double) * 3 * (size_t)(nv + 1));
  int64_t u = -1;
  for (int64_t i = 0; i < nv; ++i) {
    const double* r = V + 3 * (int64_t)idx[i];
    if (u < 0 || cmp_vrow(&idx[i], &idx[i - 1]) != 0) {
      ++u;
      memcpy(t->SV + 3 * u, r, 3 * sizeof(double));
    }
    ic[idx[i]] = (int32_t)u;
  }
  t->nsv = u + 1;
  int32_t* SF = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(nf + 1));
  for (int64_t i = 0; i < 3 * nf; ++i) SF[i] = ic[F[i]];
  free(idx);
  free(ic);
  const int64_t root = node_new(t, SF, nf);
  grow(t, root);
  return t;
}

void oracle_wntree_free(oracle_wntree* t) {
  if (!t) return;
  for (int64_t i = 0; i < t->nnodes; ++i) {
    free(t->nodes[i].F);
    free(t->nodes[i].cap);
  }
  free(t->nodes);
  free(t->SV);
  free(t);
}

int64_t oracle_wntree_num_nodes(const oracle_wntree* t) { return t->nnodes; }

/* per node, in construction (pre-)order: #faces, #cap triangles, the fan's
 * apex (an index into the de-duplicated vertices; -1 without exterior
 * edges), left and right child (-1 for a leaf) */
void oracle_wntree_nodes(const oracle_wntree* t, int64_t* out) {
  for (int64_t i = 0; i < t->nnodes; ++i) {
    const wnode* n = &t->nodes[i];
    out[5 * i] = n->nf;
    out[5 * i + 1] = n->ncap;
    out[5 * i + 2] = n->apex;
    out[5 * i + 3] = n->child[0];
    out[5 * i + 4] = n->child[1];
  }
}

static double wn_list(const oracle_wntree* t, const int32_t* T, int64_t m, const double* p) {
  double w = 0;  /* winding_number.cpp:62-89 */
  for (int64_t f = 0; f < m; ++f)
    w += oracle_solid_angle(t->SV + 3 * (int64_t)T[3 * f], t->SV + 3 * (int64_t)T[3 * f + 1],
                            t->SV + 3 * (int64_t)T[3 * f + 2], p);
  return w;
}

/* WindingNumberTree::winding_number, WindingNumberTree.h:262-336 */
static double node_wn(const oracle_wntree* t, int64_t id, const double* p) {
  const wnode* n = &t->nodes[id];
  int inside = 1; /* WindingNumberAABB::inside, closed box */
  for (int d = 0; d < 3; ++d)
    if (p[d] < n->lo[d] || p[d] > n->hi[d]) inside = 0;
  if (inside) {
    if (n->child[0] >= 0) {
      double sum = 0;
      sum += node_wn(t, n->child[0], p);
      sum += node_wn(t, n->child[1], p);
      return sum;
    }
    return wn_list(t, n->F, n->nf, p);
  }
  if ((n->ncap - 2) < n->nf) return wn_list(t, n->cap, n->ncap, p);
  return wn_list(t, n->F, n->nf, p);
}

/* solid angles node_wn evaluates for p (instrumentation for work counts) */
static int64_t node_count(const oracle_wntree* t, int64_t id, const double* p) {
  const wnode* n = &t->nodes[id];
  int inside = 1;
  for (int d = 0; d < 3; ++d)
    if (p[d] < n->lo[d] || p[d] > n->hi[d]) inside = 0;
  if (inside) {
    if (n->child[0] >= 0) return node_count(t, n->child[0], p) + node_count(t, n->child[1], p);
    return n->nf;
  }
  return (n->ncap - 2) < n->nf ? n->ncap : n->nf;
}

int64_t oracle_wntree_count(const oracle_wntree* t, const double* p) { return t->nnodes ? node_count(t, 0, p) : 0; }

typedef struct {
  const oracle_wntree* t;
  const double* O;
  double* W;
} wq_ctx;

static void wq_range(int64_t b, int64_t e, void* vctx) {
  wq_ctx* c = (wq_ctx*)vctx;
  for (int64_t o = b; o < e; ++o) c->W[o] = c->t->nnodes ? node_wn(c->t, 0, c->O + 3 * o) : 0.0;
}

/* winding_number.cpp:48-51: parallel_for over the origins, min_parallel 10000 */
void oracle_wntree_query(const oracle_wntree* t, const double* O, int64_t no, double* W, int nthreads) {
  wq_ctx c = {t, O, W};
  oracle_parallel_for(no, wq_range, &c, 10000, nthreads);
}
