// wntree.h -- the exact winding-number hierarchy (igl::WindingNumberAABB):
// host build (wntree.cpp) and its GPU records (wntree.hip).
#pragma once
#include <cstdint>
#include <vector>

namespace gwn {

// One node in depth-first pre-order ("rope" layout): its closed box, the
// record after its subtree (link), and two triangle lists in the triangle
// array: the faces a query inside a leaf sums (in_*), and what a query
// outside the box sums (out_*: the cap, or the faces when the cap is not
// smaller).  Triangles are stored as 9 doubles (three corners).
struct WNRecord {
  double lo[3], hi[3];
  uint32_t link, leaf;
  uint32_t in_start, in_count;
  uint32_t out_start, out_count;
  uint32_t pad[2];
};
static_assert(sizeof(WNRecord) == 80, "record layout");

struct WNTreeHost {
  std::vector<double> SV;        // de-duplicated vertices (lexicographic order)
  std::vector<WNRecord> rec;     // pre-order records
  std::vector<double> tri;       // triangle corners, 9 doubles each
  // per record, for parity checks: #faces, #cap triangles, fan apex (-1: no
  // exterior edges), left and right child (-1 for a leaf)
  std::vector<int64_t> nfaces, ncap, apex, left, right;
};

// Builds the reference's hierarchy of the mesh; draws rand() as the
// reference's constructors do.
void wntree_build_host(const double* V, int64_t nv, const int32_t* F, int64_t nf, WNTreeHost& T);

}  // namespace gwn
