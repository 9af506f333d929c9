/*
 * oracle_cloud_impl.h -- TEST INFRASTRUCTURE (the checker, never the product).
 * Included by oracle_cloud.c once per scalar type (REAL = double / float,
 * SFX = f64 / f32), mirroring the reference's scalar templates when every
 * Eigen argument has the same scalar type.
 *
 *   octree                      octree.cpp:7-169
 *   cloud precompute            fast_winding_number.cpp:8-124
 *   cloud evaluation            fast_winding_number.cpp:126-311
 */
#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* ---- octree ---- */
typedef struct {
  int32_t* a;
  int32_t n, cap;
} FN(ivec);

static void FN(ivec_push)(FN(ivec) * v, int32_t x) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 4;
    v->a = (int32_t*)realloc(v->a, sizeof(int32_t) * (size_t)v->cap);
  }
  v->a[v->n++] = x;
}

typedef struct {
  const REAL* P;
  FN(ivec) * pts; /* point_indices */
  int32_t* CH;    /* 8 per cell */
  REAL* CN;       /* 3 per cell */
  REAL* W;
  int32_t ncells, cap;
} FN(octb);

static int FN(oct_grow)(FN(octb) * b, int32_t need) {
  if (need <= b->cap) return 0;
  int32_t nc = b->cap ? b->cap : 64;
  while (nc < need) nc *= 2;
  b->pts = (FN(ivec)*)realloc(b->pts, sizeof(FN(ivec)) * (size_t)nc);
  b->CH = (int32_t*)realloc(b->CH, sizeof(int32_t) * 8 * (size_t)nc);
  b->CN = (REAL*)realloc(b->CN, sizeof(REAL) * 3 * (size_t)nc);
  b->W = (REAL*)realloc(b->W, sizeof(REAL) * (size_t)nc);
  if (!b->pts || !b->CH || !b->CN || !b->W) return -1;
  memset(b->pts + b->cap, 0, sizeof(FN(ivec)) * (size_t)(nc - b->cap));
  b->cap = nc;
  return 0;
}

/* helper lambda (octree.cpp:93-133): split a cell holding more than one
 * point into 8 children appended as one contiguous block, then recurse into
 * the children in octant order. */
static int FN(oct_helper)(FN(octb) * b, int32_t index, int depth) {
  if (!(b->pts[index].n > 1 && depth < 30000)) return 0;
  const int32_t m = b->ncells;
  if (FN(oct_grow)(b, m + 8)) return -1;
  for (int i = 0; i < 8; ++i) b->CH[8 * (int64_t)index + i] = m + i;
  const REAL h = b->W[index] / 2;
  const REAL c[3] = {b->CN[3 * (int64_t)index], b->CN[3 * (int64_t)index + 1], b->CN[3 * (int64_t)index + 2]};
  const REAL hh = h / 2;
  for (int i = 0; i < 8; ++i) {
    const int32_t k = m + i;
    for (int j = 0; j < 8; ++j) b->CH[8 * (int64_t)k + j] = -1;
    b->pts[k].a = NULL;
    b->pts[k].n = b->pts[k].cap = 0;
    /* translate_center (octree.cpp:55-77) */
    REAL ch[3] = {-hh, -hh, -hh};
    if (i % 2) ch[0] = hh;
    if (i == 2 || i == 3 || i == 6 || i == 7) ch[1] = hh;
    if (i > 3) ch[2] = hh;
    for (int d = 0; d < 3; ++d) b->CN[3 * (int64_t)k + d] = c[d] + ch[d];
    b->W[k] = h;
  }
  /* get_octant (octree.cpp:32-53) */
  for (int32_t j = 0; j < b->pts[index].n; ++j) {
    const int32_t p = b->pts[index].a[j];
    const REAL* x = b->P + 3 * (int64_t)p;
    int32_t o = 0;
    if (x[0] >= c[0]) o += 1;
    if (x[1] >= c[1]) o += 2;
    if (x[2] >= c[2]) o += 4;
    FN(ivec_push)(&b->pts[m + o], p);
  }
  b->ncells = m + 8;
  for (int i = 0; i < 8; ++i)
    if (FN(oct_helper)(b, m + i, depth + 1)) return -1;
  return 0;
}

int FN(oracle_octree)(const REAL* P, int32_t n, oracle_octree_t* out) {
  memset(out, 0, sizeof(*out));
  FN(octb) b;
  memset(&b, 0, sizeof(b));
  b.P = P;
  if (FN(oct_grow)(&b, 1)) return -1;
  b.pts[0].a = NULL;
  b.pts[0].n = b.pts[0].cap = 0;
  for (int32_t i = 0; i < n; ++i) FN(ivec_push)(&b.pts[0], i);
  for (int j = 0; j < 8; ++j) b.CH[j] = -1;
  /* root: bbox centre and longest bbox side (octree.cpp:139-150) */
  REAL mn[3], mx[3];
  for (int d = 0; d < 3; ++d) {
    mn[d] = n > 0 ? P[d] : 0;
    mx[d] = n > 0 ? P[d] : 0;
  }
  for (int32_t i = 1; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      const REAL v = P[3 * (int64_t)i + d];
      if (v < mn[d]) mn[d] = v;
      if (v > mx[d]) mx[d] = v;
    }
  for (int d = 0; d < 3; ++d) b.CN[d] = (mn[d] + mx[d]) / (REAL)2.0;
  REAL w = mx[0] - mn[0];
  for (int d = 1; d < 3; ++d)
    if (mx[d] - mn[d] > w) w = mx[d] - mn[d];
  b.W[0] = w;
  b.ncells = 1;
  if (FN(oct_helper)(&b, 0, 0)) return -1;

  out->ncells = b.ncells;
  out->CH = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)b.ncells);
  memcpy(out->CH, b.CH, sizeof(int32_t) * 8 * (size_t)b.ncells);
  out->CN = malloc(sizeof(REAL) * 3 * (size_t)b.ncells);
  memcpy(out->CN, b.CN, sizeof(REAL) * 3 * (size_t)b.ncells);
  out->W = malloc(sizeof(REAL) * (size_t)b.ncells);
  memcpy(out->W, b.W, sizeof(REAL) * (size_t)b.ncells);
  out->cell_start = (int64_t*)malloc(sizeof(int64_t) * (size_t)b.ncells);
  out->cell_count = (int32_t*)malloc(sizeof(int32_t) * (size_t)b.ncells);
  int64_t tot = 0;
  for (int32_t c = 0; c < b.ncells; ++c) {
    out->cell_start[c] = tot;
    out->cell_count[c] = b.pts[c].n;
    tot += b.pts[c].n;
  }
  out->npoint_list = tot;
  out->point_list = (int32_t*)malloc(sizeof(int32_t) * (size_t)(tot > 0 ? tot : 1));
  for (int32_t c = 0; c < b.ncells; ++c) {
    if (b.pts[c].n) memcpy(out->point_list + out->cell_start[c], b.pts[c].a, sizeof(int32_t) * (size_t)b.pts[c].n);
    free(b.pts[c].a);
  }
  free(b.pts);
  free(b.CH);
  free(b.CN);
  free(b.W);
  return 0;
}

/* ---- precompute (fast_winding_number.cpp:8-124) ---- */
int FN(oracle_cloud_precompute)(const REAL* P, const REAL* N, const REAL* A, const oracle_octree_t* t, int order,
                                REAL* CM, REAL* R, REAL* EC) {
  const int nt = order == 0 ? 3 : (order == 1 ? 12 : 39);
  for (int32_t index = 0; index < t->ncells; ++index) {
    REAL* ec = EC + (int64_t)nt * index;
    for (int k = 0; k < nt; ++k) ec[k] = 0;
    const int32_t* pl = t->point_list + t->cell_start[index];
    const int32_t np = t->cell_count[index];
    REAL mc[3] = {0, 0, 0}, z[3] = {0, 0, 0};
    REAL areatotal = 0;
    for (int32_t j = 0; j < np; ++j) {
      const int32_t p = pl[j];
      areatotal += A[p];
      for (int d = 0; d < 3; ++d) mc[d] += A[p] * P[3 * (int64_t)p + d];
      for (int d = 0; d < 3; ++d) z[d] += A[p] * N[3 * (int64_t)p + d];
    }
    if (np > 0) {
      for (int d = 0; d < 3; ++d) mc[d] = mc[d] / areatotal;
    } else {
      for (int d = 0; d < 3; ++d) mc[d] = (REAL)NAN;
    }
    for (int d = 0; d < 3; ++d) {
      CM[3 * (int64_t)index + d] = mc[d];
      ec[d] = z[d];
    }
    REAL max_norm = 0;
    for (int32_t j = 0; j < np; ++j) {
      const int32_t p = pl[j];
      REAL pt[3];
      for (int d = 0; d < 3; ++d) pt[d] = P[3 * (int64_t)p + d] - mc[d];
      const REAL nrm = (REAL)sqrt(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
      if (nrm > max_norm) max_norm = nrm;
      if (nt >= 12) {
        /* TempCoeffs = A * point^T * N, flattened column-major (u + 3v) */
        for (int v = 0; v < 3; ++v)
          for (int u = 0; u < 3; ++u) ec[3 + u + 3 * v] += (A[p] * pt[u]) * N[3 * (int64_t)p + v];
      }
      if (nt == 39) {
        for (int k = 0; k < 3; ++k) {
          const REAL s = (REAL)(0.5 * pt[k]);
          for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) ec[12 + 9 * k + u + 3 * v] += s * ((A[p] * pt[u]) * N[3 * (int64_t)p + v]);
        }
      }
    }
    R[index] = max_norm;
  }
  return 0;
}

/* ---- evaluation (fast_winding_number.cpp:126-311) ---- */
#define ORACLE_PI_C 3.1415926535897932384626433832795

/* direct_eval lambda (157-172): dipole loc.n / (4 pi |loc|^3); 0.5 on a point */
static inline REAL FN(direct_eval)(const REAL* loc, const REAL* an) {
  const REAL PI_4 = (REAL)(4.0 * ORACLE_PI_C);
  const REAL n = (REAL)sqrt(loc[0] * loc[0] + loc[1] * loc[1] + loc[2] * loc[2]);
  if (n == 0) return (REAL)0.5;
  return (loc[0] * an[0] + loc[1] * an[1] + loc[2] * an[2]) / (PI_4 * (n * n * n));
}

/* expansion_eval lambda (174-228): order-0/1/2 Taylor terms of the dipole
 * field about the cell's area-weighted centre. */
static REAL FN(expansion_eval)(const REAL* loc, const REAL* ec, int nt) {
  const REAL PI_4 = (REAL)(4.0 * ORACLE_PI_C);
  REAL wn = FN(direct_eval)(loc, ec);
  const REAL r = (REAL)sqrt(loc[0] * loc[0] + loc[1] * loc[1] + loc[2] * loc[2]);
  REAL PI_4_r3 = 0, PI_4_r5 = 0, PI_4_r7 = 0;
  if (nt > 3) {
    PI_4_r3 = PI_4 * r * r * r;
    PI_4_r5 = PI_4_r3 * r * r;
    const REAL d = (REAL)(1.0 / PI_4_r3);
    const REAL s = (REAL)(-3.0 / PI_4_r5);
    REAL S[9];
    for (int v = 0; v < 3; ++v)
      for (int u = 0; u < 3; ++u) S[u + 3 * v] = (loc[u] * loc[v]) * s;
    S[0] += d;
    S[4] += d;
    S[8] += d;
    REAL acc = S[0] * ec[3];
    for (int k = 1; k < 9; ++k) acc += S[k] * ec[3 + k];
    wn += acc;
  }
  if (nt > 12) {
    PI_4_r7 = PI_4_r5 * r * r;
    REAL L[9];
    for (int v = 0; v < 3; ++v)
      for (int u = 0; u < 3; ++u) L[u + 3 * v] = loc[u] * (loc[v] / PI_4_r7);
    const REAL s = (REAL)(-3.0 / PI_4_r5);
    for (int i = 0; i < 3; ++i) {
      REAL T3[9];
      for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) {
          REAL rcd = 0;
          if (u == v) rcd += loc[i];
          if (u == i) rcd += loc[v];
          if (v == i) rcd += loc[u];
          T3[u + 3 * v] = (REAL)(15.0 * loc[i]) * L[u + 3 * v] + s * rcd;
        }
      REAL acc = T3[0] * ec[12 + 9 * i];
      for (int k = 1; k < 9; ++k) acc += T3[k] * ec[12 + 9 * i + k];
      wn += acc;
    }
  }
  return wn;
}

typedef struct {
  const REAL *P, *N, *A;
  const oracle_octree_t* t;
  int nt;
  const REAL *CM, *R, *EC;
  const REAL* Q;
  double beta;
  REAL* WN;
  int32_t npts;
} FN(cev_ctx);

static REAL FN(leaf_direct)(const FN(cev_ctx) * c, int32_t cell, const REAL* q) {
  REAL wn = 0;
  const int32_t* pl = c->t->point_list + c->t->cell_start[cell];
  for (int32_t j = 0; j < c->t->cell_count[cell]; ++j) {
    const int32_t p = pl[j];
    REAL loc[3], an[3];
    for (int d = 0; d < 3; ++d) {
      loc[d] = c->P[3 * (int64_t)p + d] - q[d];
      an[d] = c->N[3 * (int64_t)p + d] * c->A[p];
    }
    wn += FN(direct_eval)(loc, an);
  }
  return wn;
}

/* helper lambda (233-292): one level of the near list; recursion adds the
 * next level's sum after this level's, i.e. l0 + (l1 + (l2 + ...)). */
static REAL FN(cev_helper)(const FN(cev_ctx) * c, const REAL* q, const int32_t* near, int32_t nnear) {
  REAL wn = 0;
  int32_t* nn = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)(nnear > 0 ? nnear : 1));
  int32_t nnn = 0;
  for (int32_t i = 0; i < nnear; ++i) {
    const int32_t index = near[i];
    if (c->t->CH[8 * (int64_t)index] == -1) {
      const int32_t* pl = c->t->point_list + c->t->cell_start[index];
      for (int32_t j = 0; j < c->t->cell_count[index]; ++j) {
        const int32_t p = pl[j];
        REAL loc[3], an[3];
        for (int d = 0; d < 3; ++d) {
          loc[d] = c->P[3 * (int64_t)p + d] - q[d];
          an[d] = c->N[3 * (int64_t)p + d] * c->A[p];
        }
        wn += FN(direct_eval)(loc, an);
      }
    } else {
      for (int ch = 0; ch < 8; ++ch) {
        const int32_t ci = c->t->CH[8 * (int64_t)index + ch];
        if (c->t->cell_count[ci] > 0) {
          REAL cmq[3];
          for (int d = 0; d < 3; ++d) cmq[d] = c->CM[3 * (int64_t)ci + d] - q[d];
          const REAL nrm = (REAL)sqrt(cmq[0] * cmq[0] + cmq[1] * cmq[1] + cmq[2] * cmq[2]);
          if ((double)nrm > c->beta * (double)c->R[ci]) {
            if (c->t->CH[8 * (int64_t)ci] == -1)
              wn += FN(leaf_direct)(c, ci, q);
            else
              wn += FN(expansion_eval)(cmq, c->EC + (int64_t)c->nt * ci, c->nt);
          } else {
            nn[nnn++] = ci;
          }
        }
      }
    }
  }
  if (nnn > 0) wn += FN(cev_helper)(c, q, nn, nnn);
  free(nn);
  return wn;
}

static void FN(cev_range)(int64_t b, int64_t e, void* vctx) {
  const FN(cev_ctx)* c = (const FN(cev_ctx)*)vctx;
  for (int64_t it = b; it < e; ++it) {
    const REAL* q = c->Q + 3 * it;
    if (c->beta > 0) {
      const int32_t root = 0;
      c->WN[it] = FN(cev_helper)(c, q, &root, 1);
    } else {
      double wn = 0;
      for (int32_t j = 0; j < c->npts; ++j) {
        REAL loc[3], an[3];
        for (int d = 0; d < 3; ++d) {
          loc[d] = c->P[3 * (int64_t)j + d] - q[d];
          an[d] = c->N[3 * (int64_t)j + d] * c->A[j];
        }
        wn += FN(direct_eval)(loc, an);
      }
      c->WN[it] = (REAL)wn;
    }
  }
}

/* Counted events of one query's evaluation (the helper's control flow, no
 * arithmetic on the moments): child-cell distance tests, direct dipole terms
 * and Taylor expansions. */
static void FN(cev_count)(const oracle_octree_t* t, const REAL* CM, const REAL* R, const REAL* q, double beta,
                          const int32_t* near, int32_t nnear, int64_t* tests, int64_t* dipoles, int64_t* exps) {
  int32_t* nn = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)(nnear > 0 ? nnear : 1));
  int32_t nnn = 0;
  for (int32_t i = 0; i < nnear; ++i) {
    const int32_t index = near[i];
    if (t->CH[8 * (int64_t)index] == -1) {
      *dipoles += t->cell_count[index];
    } else {
      for (int ch = 0; ch < 8; ++ch) {
        const int32_t ci = t->CH[8 * (int64_t)index + ch];
        if (t->cell_count[ci] > 0) {
          ++*tests;
          REAL cmq[3];
          for (int d = 0; d < 3; ++d) cmq[d] = CM[3 * (int64_t)ci + d] - q[d];
          const REAL nrm = (REAL)sqrt(cmq[0] * cmq[0] + cmq[1] * cmq[1] + cmq[2] * cmq[2]);
          if ((double)nrm > beta * (double)R[ci]) {
            if (t->CH[8 * (int64_t)ci] == -1)
              *dipoles += t->cell_count[ci];
            else
              ++*exps;
          } else {
            nn[nnn++] = ci;
          }
        }
      }
    }
  }
  if (nnn > 0) FN(cev_count)(t, CM, R, q, beta, nn, nnn, tests, dipoles, exps);
  free(nn);
}

void FN(oracle_cloud_count)(const oracle_octree_t* t, const REAL* CM, const REAL* R, const REAL* q, double beta,
                            int64_t* tests, int64_t* dipoles, int64_t* exps) {
  *tests = *dipoles = *exps = 0;
  const int32_t root = 0;
  if (t->ncells > 0) FN(cev_count)(t, CM, R, q, beta, &root, 1, tests, dipoles, exps);
}

void FN(oracle_cloud_eval)(const REAL* P, const REAL* N, const REAL* A, const oracle_octree_t* t, int order,
                           const REAL* CM, const REAL* R, const REAL* EC, const REAL* Q, int64_t nq, double beta,
                           REAL* WN, int nthreads) {
  FN(cev_ctx) c;
  c.P = P;
  c.N = N;
  c.A = A;
  c.t = t;
  c.nt = order == 0 ? 3 : (order == 1 ? 12 : 39);
  c.CM = CM;
  c.R = R;
  c.EC = EC;
  c.Q = Q;
  c.beta = beta;
  c.WN = WN;
  c.npts = t->ncells > 0 ? t->cell_count[0] : 0;
  oracle_parallel_for(nq, FN(cev_range), &c, 1000, nthreads);
}

#undef FN
#undef CAT
#undef CAT_
