/*
 * oracle_cloud.c -- TEST INFRASTRUCTURE (the checker, never the product).
 * Point-cloud octree, precompute and evaluation restated from octree.cpp and
 * fast_winding_number.cpp:8-311, instantiated for double and float.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define REAL double
#define SFX f64
#include "oracle_cloud_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX f32
#include "oracle_cloud_impl.h"
#undef REAL
#undef SFX

void oracle_octree_release(oracle_octree_t* t) {
  if (!t) return;
  free(t->CH);
  free(t->CN);
  free(t->W);
  free(t->cell_start);
  free(t->cell_count);
  free(t->point_list);
  memset(t, 0, sizeof(*t));
}
