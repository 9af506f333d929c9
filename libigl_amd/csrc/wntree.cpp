// wntree.cpp -- host build of the exact winding-number hierarchy.
//
// igl::winding_number(V,F,O,W) (winding_number.cpp:39-53) builds an
// igl::WindingNumberAABB (WindingNumberAABB.h:149-235 over
// WindingNumberTree.h:150-336) and sums, per query, the solid angles of a
// node's faces or, for a node whose box does not contain the query, of its
// cap: the triangle fan that closes the node's open boundary.  Outside the
// box the closed surface (faces + reversed cap) has winding number 0, so both
// sums are equal in exact arithmetic; caps have far fewer triangles.
//
// The tree is the reference's: duplicate vertices merged (unique rows in
// lexicographic order, remove_duplicate_vertices.cpp / unique_rows.cpp);
// each node's exterior edges (exterior_edges.cpp:46-100) fanned around the
// first vertex of the edge rand() % #edges picks (triangle_fan.cpp:15-47);
// a node with more than 100 faces and a cap smaller than its faces
// (#cap - 2 < #F) splits at the median (median.cpp: std::nth_element at n,
// then at n - 1 for an even count, m = (v[n] + v[n-1]) / 2) of its face
// barycenters (barycenter.cpp) along the longest side of its box, faces
// <= median going left in their order; nodes are constructed, and draw
// rand(), in depth-first pre-order.
//
// Built in two passes.  (1) Every node with more than 100 faces is split and
// the exterior edges of every node are found, independently of rand(); large
// subtrees run on their own threads.  (2) One pre-order pass draws rand() for
// each node that has exterior edges, exactly where the reference's
// constructors would, fixes its cap, and drops the children of nodes whose
// cap makes them leaves.  The result is flattened into a rope of records for
// the GPU walk (wntree.hip).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <future>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "gwn_internal.h"
#include "wntree.h"

namespace gwn {
namespace {

constexpr int64_t kLeafFaces = 100;          // WindingNumberAABB_MIN_F
constexpr int64_t kParallelFaces = 1 << 14;  // subtrees at least this big get their own thread

struct Node {
  int64_t b = 0, e = 0;  // faces perm[b, e), in the reference's order
  double lo[3], hi[3];
  std::vector<int32_t> E;  // exterior edges, 2 per row, in exterior_edges' order
  std::unique_ptr<Node> child[2];
  // fixed by the rand() pass
  int32_t apex = -1;
  int64_t ncap = 0;
  bool leaf = true;
};

struct Build {
  const double* SV = nullptr;
  const int32_t* SF = nullptr;
  std::vector<int32_t> perm;  // face ids, partitioned stably per node
};

// Radix sort of (key, tag) pairs by key, least significant byte first, over
// the bytes in which the keys differ.
void radix_pairs(std::vector<uint64_t>& key, std::vector<int8_t>& tag) {
  const size_t n = key.size();
  if (n < 2) return;
  uint64_t anyd = 0;
  for (size_t i = 1; i < n; ++i) anyd |= key[i] ^ key[0];
  std::vector<uint64_t> k2(n);
  std::vector<int8_t> t2(n);
  for (int shift = 0; shift < 64; shift += 8) {
    if (((anyd >> shift) & 0xFF) == 0) continue;
    size_t cnt[257] = {0};
    for (size_t i = 0; i < n; ++i) ++cnt[((key[i] >> shift) & 0xFF) + 1];
    for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
    for (size_t i = 0; i < n; ++i) {
      const size_t p = cnt[(key[i] >> shift) & 0xFF]++;
      k2[p] = key[i];
      t2[p] = tag[i];
    }
    key.swap(k2);
    tag.swap(t2);
  }
}

// exterior_edges.cpp:46-100 over faces perm[b, e): each facet edge
// (oriented_facets.cpp:36-43: (1,2), then (2,0), then (0,1)) counted +1 when
// it runs from the lower index to the higher (sort2 swaps only on a > b), -1
// otherwise; unique edges in ascending (min, max) order; each non-zero edge
// repeated |count| times, from min to max for a positive count.
void exterior_edges(const Build& B, int64_t b, int64_t e, std::vector<int32_t>& E) {
  const int64_t m = e - b;
  std::vector<uint64_t> key(3 * (size_t)m);
  std::vector<int8_t> tag(3 * (size_t)m);
  for (int c = 0; c < 3; ++c)
    for (int64_t i = 0; i < m; ++i) {
      const int32_t* f = B.SF + 3 * (int64_t)B.perm[b + i];
      const int32_t u = f[(c + 1) % 3], v = f[(c + 2) % 3];
      const bool up = !(u > v);
      const uint64_t lo = (uint32_t)(up ? u : v), hi = (uint32_t)(up ? v : u);
      key[c * m + i] = (lo << 32) | hi;
      tag[c * m + i] = up ? 1 : -1;
    }
  radix_pairs(key, tag);
  E.clear();
  for (size_t a = 0; a < key.size();) {
    size_t z = a;
    int64_t count = 0;
    while (z < key.size() && key[z] == key[a]) count += tag[z++];
    const int32_t lo = (int32_t)(key[a] >> 32), hi = (int32_t)(key[a] & 0xFFFFFFFFu);
    for (int64_t r = 0; r < (count < 0 ? -count : count); ++r) {
      E.push_back(count > 0 ? lo : hi);
      E.push_back(count > 0 ? hi : lo);
    }
    a = z;
  }
}

// compute_min_max_corners (WindingNumberAABB.h:255-294)
void corners(const Build& B, Node& n) {
  for (int d = 0; d < 3; ++d) {
    n.lo[d] = std::numeric_limits<double>::infinity();
    n.hi[d] = -std::numeric_limits<double>::infinity();
  }
  for (int64_t i = n.b; i < n.e; ++i)
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) {
        const double v = B.SV[3 * (int64_t)B.SF[3 * (int64_t)B.perm[i] + j] + d];
        n.lo[d] = v < n.lo[d] ? v : n.lo[d];
        n.hi[d] = v > n.hi[d] ? v : n.hi[d];
      }
}

// Pass 1: the node's edges and box, and (above the face limit) its split,
// recursively.  The cap rule may later turn the node into a leaf; the split
// itself does not depend on it.
void speculate(Build& B, Node& n) {
  exterior_edges(B, n.b, n.e, n.E);
  corners(B, n);
  const int64_t nf = n.e - n.b;
  if (nf <= kLeafFaces) return;
  int max_d = -1;
  double max_len = -std::numeric_limits<double>::infinity();
  for (int d = 0; d < 3; ++d)
    if (n.hi[d] - n.lo[d] > max_len) {
      max_len = n.hi[d] - n.lo[d];
      max_d = d;
    }
  std::vector<double> bc((size_t)nf);
  for (int64_t i = 0; i < nf; ++i) {
    const int32_t* f = B.SF + 3 * (int64_t)B.perm[n.b + i];
    double s = 0.0;  // barycenter.cpp: zero, add the three corners, divide by 3
    for (int j = 0; j < 3; ++j) s += B.SV[3 * (int64_t)f[j] + max_d];
    bc[i] = s / 3.0;
  }
  std::vector<double> v(bc);
  const size_t h = (size_t)nf / 2;
  std::nth_element(v.begin(), v.begin() + h, v.end());
  double split;
  if (nf % 2 == 0) {
    std::nth_element(v.begin(), v.begin() + h - 1, v.end());
    split = 0.5 * (v[h] + v[h - 1]);
  } else {
    split = v[h];
  }
  std::vector<int32_t> right;
  right.reserve((size_t)nf);
  int64_t w = n.b;
  for (int64_t i = 0; i < nf; ++i) {
    const int32_t f = B.perm[n.b + i];
    if (bc[i] <= split)
      B.perm[w++] = f;
    else
      right.push_back(f);
  }
  const int64_t nl = w - n.b;
  std::copy(right.begin(), right.end(), B.perm.begin() + w);
  if (nl == 0 || nl == nf) return;  // one side empty: a leaf (WindingNumberAABB.h:196-201)
  for (int c = 0; c < 2; ++c) n.child[c].reset(new Node());
  n.child[0]->b = n.b;
  n.child[0]->e = w;
  n.child[1]->b = w;
  n.child[1]->e = n.e;
  bc.clear();
  bc.shrink_to_fit();
  v.clear();
  v.shrink_to_fit();
  if (nl >= kParallelFaces && nf - nl >= kParallelFaces && std::thread::hardware_concurrency() > 1) {
    auto left = std::async(std::launch::async, [&B, &n] { speculate(B, *n.child[0]); });
    speculate(B, *n.child[1]);
    left.get();
  } else {
    speculate(B, *n.child[0]);
    speculate(B, *n.child[1]);
  }
}

// Pass 2: the rand() draws in construction order (a node's constructor,
// then its left subtree, then its right subtree) and the cap rule of grow()
// (WindingNumberAABB.h:154-160).
void draw(Node& n) {
  const int64_t ne = (int64_t)n.E.size() / 2;
  n.ncap = 0;
  n.apex = -1;
  if (ne > 0) {
    n.apex = n.E[2 * ((int64_t)std::rand() % ne)];
    for (int64_t i = 0; i < ne; ++i) n.ncap += (n.E[2 * i] != n.apex && n.E[2 * i + 1] != n.apex);
  }
  const int64_t nf = n.e - n.b;
  n.leaf = !n.child[0] || nf <= kLeafFaces || (n.ncap - 2) >= nf;
  if (n.leaf) {
    n.child[0].reset();
    n.child[1].reset();
    return;
  }
  draw(*n.child[0]);
  draw(*n.child[1]);
}

void put_tri(const Build& B, int32_t a, int32_t b, int32_t c, std::vector<double>& tri) {
  for (int32_t v : {a, b, c})
    for (int d = 0; d < 3; ++d) tri.push_back(B.SV[3 * (int64_t)v + d]);
}

// Pass 3: pre-order records.  Outside a node's box the walk sums its cap
// while #cap - 2 < #F, else its faces; inside a leaf, its faces.  Returns the
// node's record index.
int64_t flatten(const Build& B, const Node& n, WNTreeHost& T) {
  const int64_t id = (int64_t)T.rec.size();
  const int64_t nf = n.e - n.b;
  T.rec.emplace_back();
  T.nfaces.push_back(nf);
  T.ncap.push_back(n.ncap);
  T.apex.push_back(n.apex);
  T.left.push_back(-1);
  T.right.push_back(-1);
  WNRecord r{};
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = n.lo[d];
    r.hi[d] = n.hi[d];
  }
  auto faces = [&]() {
    const uint32_t start = (uint32_t)(T.tri.size() / 9);
    for (int64_t i = n.b; i < n.e; ++i) {
      const int32_t* f = B.SF + 3 * (int64_t)B.perm[i];
      put_tri(B, f[0], f[1], f[2], T.tri);
    }
    return start;
  };
  if (n.leaf) {
    r.in_start = faces();
    r.in_count = (uint32_t)nf;
  }
  if ((n.ncap - 2) < nf) {
    r.out_start = (uint32_t)(T.tri.size() / 9);
    r.out_count = (uint32_t)n.ncap;
    const int64_t ne = (int64_t)n.E.size() / 2;
    for (int64_t i = 0; i < ne; ++i) {
      const int32_t a = n.E[2 * i], b = n.E[2 * i + 1];
      if (a == n.apex || b == n.apex) continue;
      put_tri(B, n.apex, a, b, T.tri);
    }
  } else if (n.leaf) {
    r.out_start = r.in_start;
    r.out_count = r.in_count;
  } else {
    r.out_start = faces();
    r.out_count = (uint32_t)nf;
  }
  r.leaf = n.leaf ? 1u : 0u;
  if (!n.leaf) {
    T.left[id] = flatten(B, *n.child[0], T);
    T.right[id] = flatten(B, *n.child[1], T);
  }
  r.link = (uint32_t)T.rec.size();
  T.rec[id] = r;
  return id;
}

}  // namespace

void wntree_build_host(const double* V, int64_t nv, const int32_t* F, int64_t nf, WNTreeHost& T) {
  GWN_CHECK(nv < (int64_t)0x7FFFFFFF && nf < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "mesh too large");
  // remove_duplicate_vertices(V,F,0): unique rows, sorted lexicographically
  std::vector<int32_t> idx((size_t)nv);
  for (int64_t i = 0; i < nv; ++i) idx[i] = (int32_t)i;
  auto less = [V](int32_t a, int32_t b) {
    for (int c = 0; c < 3; ++c) {
      if (V[3 * (int64_t)a + c] < V[3 * (int64_t)b + c]) return true;
      if (V[3 * (int64_t)b + c] < V[3 * (int64_t)a + c]) return false;
    }
    return false;
  };
  std::sort(idx.begin(), idx.end(), less);
  std::vector<int32_t> ic((size_t)nv);
  T.SV.clear();
  int64_t u = -1;
  for (int64_t i = 0; i < nv; ++i) {
    if (u < 0 || less(idx[i - 1], idx[i])) {
      ++u;
      for (int c = 0; c < 3; ++c) T.SV.push_back(V[3 * (int64_t)idx[i] + c]);
    }
    ic[idx[i]] = (int32_t)u;
  }
  std::vector<int32_t> SF(3 * (size_t)nf);
  for (int64_t i = 0; i < 3 * nf; ++i) SF[i] = ic[F[i]];
  Build B;
  B.SV = T.SV.data();
  B.SF = SF.data();
  B.perm.resize((size_t)nf);
  for (int64_t i = 0; i < nf; ++i) B.perm[i] = (int32_t)i;
  Node root;
  root.b = 0;
  root.e = nf;
  speculate(B, root);
  draw(root);
  T.rec.clear();
  T.tri.clear();
  T.nfaces.clear();
  T.ncap.clear();
  T.apex.clear();
  T.left.clear();
  T.right.clear();
  flatten(B, root, T);
}

}  // namespace gwn
