// sah4_build.cpp -- the reference tree topology, built on the host.
//
// Parity mode of the soup path: the GPU computes moments and traverses, but
// the 4-wide tree itself must be the one the reference builds, or the Taylor
// truncation error (~3e-3 at accuracy_scale 2) would differ from the
// reference's.  This file follows UT::BVH<4>::init<BOX_AREA>
// (FastWindingNumberForSoups.h:4579-4634) step for step: binned SAH over 16
// spans with at least 1/16 of the boxes per side (5580-5799), sorted sweep
// for <= 32 boxes (5467-5578), exhaustive partitions for <= 6 boxes
// (5362-5451), the 4-way multiSplit that splits the child of largest
// area x count (5296-5346), the median fallback (5708-5738, nthElement
// 5845-5898, partitionByCentre 5900-5944) and the node layout in which each
// child node precedes its own subtree (4916-5047).  Floating-point steps keep
// the reference's operation order (the file is built with -ffp-contract=off).
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <future>
#include <limits>
#include <thread>
#include <vector>

#include "gwn_internal.h"

namespace gwn {
namespace {

constexpr uint32_t kN = 4;
constexpr uint32_t kSpans = 16;
constexpr uint32_t kSplits = kSpans - 1;
constexpr uint32_t kMinFraction = 16;
constexpr uint32_t kSmall = 6;
constexpr uint32_t kMid = 2 * kSpans;

struct Box {
  float lo[3], hi[3];
  static Box empty() {
    Box b;
    for (int a = 0; a < 3; ++a) {
      b.lo[a] = FLT_MAX;
      b.hi[a] = -FLT_MAX;
    }
    return b;
  }
  void grow(const Box& o) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = (lo[a] < o.lo[a]) ? lo[a] : o.lo[a];
      hi[a] = (hi[a] > o.hi[a]) ? hi[a] : o.hi[a];
    }
  }
  // Box::half_surface_area, NAXES == 2 (4071-4075) or 3 (4076-4080)
  float half_area(int naxes) const {
    const float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1];
    if (naxes == 2) return d0 + d1;
    const float d2 = hi[2] - lo[2];
    return d0 * d1 + d1 * d2 + d2 * d0;
  }
  float centre2(uint32_t a) const { return lo[a] + hi[a]; }
  bool non_finite() const {
    for (int a = 0; a < 3; ++a) {
      uint32_t u, v;
      std::memcpy(&u, &lo[a], 4);
      std::memcpy(&v, &hi[a], 4);
      if ((u & 0x7F800000u) == 0x7F800000u || (v & 0x7F800000u) == 0x7F800000u) return true;
    }
    return false;
  }
};

// Subtrees of at least this many boxes are built on their own threads when a
// node has two or more of them (the reference's initNode does the same,
// 4942-5031); the node array comes out as the serial order: every child node
// precedes its own subtree, children in slot order.
constexpr uint32_t kParallelBoxes = 1024;

class Builder {
 public:
  Builder(const std::vector<Box>& boxes, std::vector<uint32_t>& out, int naxes)
      : boxes_(boxes), out_(out), naxes_(naxes),
        max_threads_(std::min(32u, std::max(1u, std::thread::hardware_concurrency()))) {}

  void run(std::vector<uint32_t>& idx) {
    const uint32_t n = (uint32_t)idx.size();
    Box full = boxes_[idx[0]];
    for (uint32_t i = 1; i < n; ++i) full.grow(boxes_[idx[i]]);
    // The root bound is computeFullBoundingBox (4856-4913) over the kept
    // boxes.  From 8192 boxes on, on a host with more than one hardware
    // thread, its parallel branch folds the boxes into per-thread boxes that
    // UT_Array::setSize zero-fills (a POD Box, 2342-2360), so the root bound
    // also contains the origin.  The root split depends on it.
    if (n >= 2 * 4096 && std::thread::hardware_concurrency() > 1) {
      for (int a = 0; a < 3; ++a) {
        full.lo[a] = (0.0f < full.lo[a]) ? 0.0f : full.lo[a];
        full.hi[a] = (0.0f > full.hi[a]) ? 0.0f : full.hi[a];
      }
    }
    out_.assign(4, NODE_EMPTY);
    build_node(out_, 0, full, idx.data(), n);
  }

 private:
  const std::vector<Box>& boxes_;
  std::vector<uint32_t>& out_;
  const int naxes_;
  const unsigned max_threads_;
  std::atomic<unsigned> threads_{1};

  float c2(uint32_t i, uint32_t axis) const { return boxes_[i].centre2(axis); }

  // Three-way partition about pivotx2: [< pivot | == pivot | > pivot], with the
  // reference's swap pattern so that the resulting order matches.
  void partition(uint32_t* b, uint32_t* e, uint32_t axis, float pivotx2, uint32_t*& ps_out,
                 uint32_t*& pe_out) const {
    uint32_t* lt = b;  // first element >= pivot
    uint32_t* le = b;  // first element > pivot
    for (uint32_t* p = b; p != e; ++p) {
      const float s = c2(*p, axis);
      if (s < pivotx2) {
        if (p != lt) {
          const uint32_t t = *p;
          if (lt == le) {
            *p = *lt;
          } else {
            *p = *le;
            *le = *lt;
          }
          *lt = t;
        }
        ++lt;
        ++le;
      } else if (s == pivotx2) {
        if (p != le) std::swap(*p, *le);
        ++le;
      }
    }
    ps_out = lt;
    pe_out = le;
  }

  void select_nth(uint32_t* b, uint32_t* e, uint32_t axis, uint32_t* nth) const {
    for (;;) {
      float a = c2(b[0], axis), m = c2(b[(e - b) / 2], axis), z = c2(*(e - 1), axis);
      // order the three candidates descending and take the middle one
      if (a < m) std::swap(a, m);
      if (a < z) std::swap(a, z);
      if (m < z) std::swap(m, z);
      uint32_t *ps, *pe;
      partition(b, e, axis, m, ps, pe);
      if (nth < ps)
        e = ps;
      else if (nth < pe)
        return;
      else
        b = pe;
      if (e <= b + 1) return;
    }
  }

  Box union_of(const uint32_t* b, const uint32_t* e) const {
    Box r = boxes_[*b];
    for (const uint32_t* p = b + 1; p < e; ++p) r.grow(boxes_[*p]);
    return r;
  }

  // Two-way split of [idx, idx+n); returns the split point, fills halves[2].
  uint32_t* split(const Box bound, uint32_t* idx, uint32_t n, Box halves[2]) const {
    if (n == 2) {
      halves[0] = boxes_[idx[0]];
      halves[1] = boxes_[idx[1]];
      return idx + 1;
    }
    if (n <= kSmall) return split_small(idx, n, halves);

    uint32_t axis = 0;
    float len = bound.hi[0] - bound.lo[0];
    for (uint32_t a = 1; a < (uint32_t)naxes_; ++a) {
      const float l = bound.hi[a] - bound.lo[a];
      if (l > len) {
        axis = a;
        len = l;
      }
    }
    if (!(len > 0.0f)) {
      halves[0] = bound;
      halves[1] = bound;
      return idx + n / 2;
    }
    if (n <= kMid) return split_sweep(idx, n, axis, halves);
    return split_binned(bound, idx, n, axis, len, halves);
  }

  uint32_t* split_small(uint32_t* idx, uint32_t n, Box halves[2]) const {
    Box local[kSmall];
    for (uint32_t i = 0; i < n; ++i) local[i] = boxes_[idx[i]];
    const uint32_t limit = 1u << (n - 1);
    uint32_t best = 0;
    float best_cost = 0.0f;
    for (uint32_t bits = 1; bits < limit; ++bits) {
      Box part[2] = {local[0], Box::empty()};
      uint32_t count[2] = {1, 0};
      for (uint32_t k = 0; k + 1 < n; ++k) {
        const uint32_t side = (bits >> k) & 1u;
        part[side].grow(local[k + 1]);
        ++count[side];
      }
      const float cost = part[0].half_area(naxes_) * (float)count[0] + part[1].half_area(naxes_) * (float)count[1];
      if (best == 0 || cost < best_cost) {
        best = bits;
        best_cost = cost;
        halves[0] = part[0];
        halves[1] = part[1];
      }
    }
    uint32_t rest[kSmall - 1];
    std::copy(idx + 1, idx + n, rest);
    uint32_t* w = idx + 1;
    for (uint32_t k = 0; k + 1 < n; ++k)
      if (!((best >> k) & 1u)) *w++ = rest[k];
    uint32_t* mid = w;
    for (uint32_t k = 0; k + 1 < n; ++k)
      if ((best >> k) & 1u) *w++ = rest[k];
    return mid;
  }

  uint32_t* split_sweep(uint32_t* idx, uint32_t n, uint32_t axis, Box halves[2]) const {
    // stable order by doubled centre along the axis
    std::stable_sort(idx, idx + n, [&](uint32_t a, uint32_t b) { return c2(a, axis) < c2(b, axis); });
    const uint32_t ns = n - 1;
    Box left[kMid - 1], right[kMid - 1];
    Box acc = boxes_[idx[0]];
    left[0] = acc;
    for (uint32_t i = 1; i < ns; ++i) {
      acc.grow(boxes_[idx[i]]);
      left[i] = acc;
    }
    // The reference seeds the right-hand sweep with idx[ns-1] and walks down to
    // idx[1], so the last box never enters a right-hand bound; kept as is
    // because these bounds steer the children's later splits.
    acc = boxes_[idx[ns - 1]];
    right[ns - 1] = acc;
    for (uint32_t i = ns - 1; i > 0; --i) {
      acc.grow(boxes_[idx[i]]);
      right[i - 1] = acc;
    }
    uint32_t best = 0;
    float best_cost = left[0].half_area(naxes_) + right[0].half_area(naxes_) * (float)(n - 1);
    for (uint32_t s = 1; s < ns; ++s) {
      const float cost = left[s].half_area(naxes_) * (float)(s + 1) + right[s].half_area(naxes_) * (float)(n - (s + 1));
      if (cost < best_cost) {
        best = s;
        best_cost = cost;
      }
    }
    halves[0] = left[best];
    halves[1] = right[best];
    return idx + best + 1;
  }

  uint32_t* split_binned(const Box& bound, uint32_t* idx, uint32_t n, uint32_t axis, float len, Box halves[2]) const {
    Box span_box[kSpans];
    uint32_t span_count[kSpans];
    for (uint32_t i = 0; i < kSpans; ++i) {
      span_box[i] = Box::empty();
      span_count[i] = 0;
    }
    const float lo2 = 2.0f * bound.lo[axis];
    const float scale = (0.5f * (float)kSpans) / len;
    for (uint32_t k = 0; k < n; ++k) {
      const Box& b = boxes_[idx[k]];
      int s = (int)((b.centre2(axis) - lo2) * scale);
      s = (s <= 0) ? 0 : ((s >= (int)kSpans - 1) ? (int)kSpans - 1 : s);
      ++span_count[s];
      span_box[s].grow(b);
    }
    Box left[kSplits], right[kSplits];
    uint32_t left_count[kSplits];
    Box acc = span_box[0];
    left[0] = acc;
    for (uint32_t i = 1; i < kSplits; ++i) {
      acc.grow(span_box[i]);
      left[i] = acc;
    }
    acc = span_box[kSpans - 1];
    right[kSplits - 1] = acc;
    for (uint32_t i = kSplits - 1; i > 0; --i) {
      acc.grow(span_box[i]);
      right[i - 1] = acc;
    }
    uint32_t run = span_count[0];
    left_count[0] = run;
    for (uint32_t i = 1; i < kSplits; ++i) {
      run += span_count[i];
      left_count[i] = run;
    }
    const uint32_t min_count = n / kMinFraction;
    const uint32_t max_count = (uint32_t)(((uint64_t)(kMinFraction - 1) * n) / kMinFraction);
    float best_cost = std::numeric_limits<float>::infinity();
    int best = -1;
    for (uint32_t s = 0; s < kSplits; ++s) {
      const uint32_t lc = left_count[s];
      if (lc < min_count || lc > max_count) continue;
      const float cost = (float)lc * left[s].half_area(naxes_) + (float)(n - lc) * right[s].half_area(naxes_);
      if (cost < best_cost) {
        best_cost = cost;
        best = (int)s;
      }
    }
    uint32_t* end = idx + n;
    if (best < 0) {
      uint32_t* nth;
      if (left_count[0] > max_count)
        nth = idx + max_count;
      else if (left_count[kSplits - 1] < min_count)
        nth = idx + min_count;
      else
        nth = idx + n / 2;
      select_nth(idx, end, axis, nth);
      halves[0] = union_of(idx, nth);
      halves[1] = union_of(nth, end);
      return nth;
    }
    const float pivot = lo2 + ((float)(best + 1) * len) / (float)(kSpans / 2);
    uint32_t *ps, *pe;
    partition(idx, end, axis, pivot, ps, pe);
    uint32_t* mid = idx + left_count[best];
    if (mid >= ps && mid <= pe) {
      halves[0] = left[best];
      halves[1] = right[best];
      return mid;
    }
    // rounding moved the partition: clamp to it and recompute the bounds
    mid = (mid < ps) ? ps : pe;
    if (mid == idx)
      ++mid;
    else if (mid == end)
      --mid;
    halves[0] = union_of(idx, mid);
    halves[1] = union_of(mid, end);
    return mid;
  }

  // 4-way split: split in two, then twice more split the part with the
  // largest area x count among those with more than one box.
  void split4(const Box& bound, uint32_t* idx, uint32_t n, uint32_t* part[kN + 1], Box part_box[kN]) const {
    part[0] = idx;
    part[2] = idx + n;
    part[1] = split(bound, idx, n, part_box);
    float area[kN] = {part_box[0].half_area(naxes_), part_box[1].half_area(naxes_), 0, 0};
    for (uint32_t np = 2; np < kN; ++np) {
      int pick = -1;
      float pick_cost = 0.0f;
      for (uint32_t i = 0; i < np; ++i) {
        const uint32_t cnt = (uint32_t)(part[i + 1] - part[i]);
        if (cnt > 1) {
          const float cost = area[i] * (float)cnt;
          if (pick < 0 || cost > pick_cost) {
            pick = (int)i;
            pick_cost = cost;
          }
        }
      }
      const uint32_t c = (uint32_t)pick;
      uint32_t* b = part[c];
      uint32_t* e = part[c + 1];
      for (uint32_t i = np; i > c; --i) part[i + 1] = part[i];
      for (uint32_t i = np - 1; i > c; --i) {
        part_box[i + 1] = part_box[i];
        area[i + 1] = area[i];
      }
      const Box parent = part_box[c];
      part[c + 1] = split(parent, b, (uint32_t)(e - b), &part_box[c]);
      area[c] = part_box[c].half_area(naxes_);
      area[c + 1] = part_box[c + 1].half_area(naxes_);
    }
  }

  // Builds node `node` of `out` over idx[0, n) and appends its subtree.
  void build_node(std::vector<uint32_t>& out, uint32_t node, const Box& bound, uint32_t* idx, uint32_t n) {
    if (n <= kN) {
      for (uint32_t i = 0; i < kN; ++i) out[4 * node + i] = (i < n) ? idx[i] : NODE_EMPTY;
      return;
    }
    uint32_t* part[kN + 1];
    Box part_box[kN];
    split4(bound, idx, n, part, part_box);
    for (uint32_t i = 0; i < kN; ++i)
      if (part[i + 1] - part[i] == 1) out[4 * node + i] = part[i][0];
    // large children on threads of their own, each into a local array whose
    // node 0 is the child; spliced in below at the child's serial position
    std::vector<uint32_t> local[kN];
    std::future<void> task[kN];
    int nlarge = 0;
    for (uint32_t i = 0; i < kN; ++i) nlarge += (uint32_t)(part[i + 1] - part[i]) >= kParallelBoxes;
    if (nlarge >= 2) {
      for (uint32_t i = 0; i < kN; ++i) {
        const uint32_t cnt = (uint32_t)(part[i + 1] - part[i]);
        if (cnt < kParallelBoxes) continue;
        local[i].reserve((size_t)4 * (cnt / 2 + cnt / 6 + 4));
        local[i].assign(4, NODE_EMPTY);
        if (threads_.fetch_add(1) < max_threads_) {
          task[i] = std::async(std::launch::async, [this, &local, &part_box, &part, i, cnt] {
            build_node(local[i], 0, part_box[i], part[i], cnt);
            threads_.fetch_sub(1);
          });
        } else {
          threads_.fetch_sub(1);
          build_node(local[i], 0, part_box[i], part[i], cnt);
        }
      }
    }
    for (uint32_t i = 0; i < kN; ++i) {
      const uint32_t cnt = (uint32_t)(part[i + 1] - part[i]);
      if (cnt == 1) continue;
      const uint32_t child = (uint32_t)(out.size() / 4);
      out[4 * node + i] = child | NODE_INTERNAL;
      if (!local[i].empty()) {
        if (task[i].valid()) task[i].get();
        const size_t at = out.size();
        out.insert(out.end(), local[i].begin(), local[i].end());
        for (size_t k = at; k < out.size(); ++k)
          if (out[k] != NODE_EMPTY && (out[k] & NODE_INTERNAL)) out[k] += child;
        std::vector<uint32_t>().swap(local[i]);
      } else {
        out.insert(out.end(), 4, NODE_EMPTY);
        build_node(out, child, part_box[i], part[i], cnt);
      }
    }
  }
};

}  // namespace

void sah4_build(const float* boxes6, uint32_t nboxes, std::vector<uint32_t>& nodes, int naxes) {
  nodes.clear();
  if (nboxes == 0) return;
  std::vector<Box> boxes(nboxes);
  for (uint32_t i = 0; i < nboxes; ++i)
    for (int a = 0; a < 3; ++a) {
      boxes[i].lo[a] = boxes6[6 * (size_t)i + 2 * a];
      boxes[i].hi[a] = boxes6[6 * (size_t)i + 2 * a + 1];
    }
  // drop boxes with NaN/inf coordinates, keeping the order of the rest
  // (utExcludeNaNInfBoxIndices, 4535-4577)
  std::vector<uint32_t> idx;
  idx.reserve(nboxes);
  for (uint32_t i = 0; i < nboxes; ++i)
    if (!boxes[i].non_finite()) idx.push_back(i);
  if (idx.empty()) return;
  nodes.reserve((size_t)4 * (nboxes / 2 + nboxes / 6 + 4));
  Builder(boxes, nodes, naxes).run(idx);
}

}  // namespace gwn
