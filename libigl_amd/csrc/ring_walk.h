// ring_walk.h -- one query's walk of a rope-ordered tree shared by the 64
// lanes of a wavefront.
//
// A sequential walk is a chain of dependent record fetches as long as the
// query's path.  Here the wavefront keeps a ring of pending (first record,
// end) ranges of sibling records in LDS; each round every lane takes one
// range, evaluates its first record, and pushes the record's child range
// [i+1, link) if the walk descends there and the rest of the range (from i+1
// after a leaf, from link after an internal record) if not empty, so a round
// needs no fetch beyond the record itself.  When the ring is full a lane walks
// its ranges sequentially instead.  Every record the sequential walk would
// decide on is decided on here too, each by one lane.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gwn {

constexpr int kReevalRing = 512;            // ranges per wavefront (4 KiB of LDS)
constexpr uint32_t kRopeLeaf = 0x80000000u;  // link word of a leaf record

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// LDS writes of the wavefront's lanes become visible to all of them
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The ring walk of one query by the calling wavefront (ring: its kReevalRing
// entries of LDS).  visit(i, link) evaluates record i for the query, adds its
// term (if any) to the lane's own partial sums, sets link to the record's link
// word and returns false when the walk descends into the record.
template <typename Visit>
__device__ __forceinline__ void ring_walk(uint2* __restrict__ ring, int lane, uint32_t nslots, Visit&& visit) {
  if (lane == 0) ring[0] = make_uint2(0u, nslots);
  uint32_t head = 0, tail = 1;  // wave-uniform ring positions
  wave_lds_sync();
  while (head != tail) {
    const uint32_t take = min(tail - head, 64u);
    const bool have = (uint32_t)lane < take;
    uint2 it = make_uint2(0u, 0u);
    if (have) it = ring[(head + lane) % kReevalRing];
    head += take;
    uint2 push0 = make_uint2(0u, 0u), push1 = make_uint2(0u, 0u);
    uint32_t np = 0;
    if (have) {
      const uint32_t i = it.x;
      uint32_t link;
      const bool done = visit(i, link);
      const uint32_t next = (link & kRopeLeaf) ? i + 1 : link;
      if (!done) push0 = make_uint2(i + 1, link), np = 1;  // child range
      if (next < it.y) {
        if (np) push1 = make_uint2(next, it.y);
        else push0 = make_uint2(next, it.y);
        ++np;
      }
    }
    const uint64_t b0 = __ballot(np & 1u), b1 = __ballot(np >> 1);
    const uint32_t below = lanes_below(b0) + 2u * lanes_below(b1);
    const uint32_t room = kReevalRing - (tail - head);
    const bool fits = below + np <= room;
    if (fits) {
      if (np > 0) ring[(tail + below) % kReevalRing] = push0;
      if (np > 1) ring[(tail + below + 1) % kReevalRing] = push1;
    } else {
      // ring full: walk the ranges sequentially (rope order)
      for (uint32_t m = 0; m < np; ++m) {
        const uint2 rg = m ? push1 : push0;
        uint32_t i = rg.x;
        while (i < rg.y) {
          uint32_t link;
          const bool done = visit(i, link);
          i = (done && !(link & kRopeLeaf)) ? link : i + 1;
        }
      }
    }
    // lanes that fit are a prefix in lane order: their ranges are contiguous
    tail += (uint32_t)__popcll(__ballot(fits && (np & 1u))) + 2u * (uint32_t)__popcll(__ballot(fits && (np >> 1)));
    wave_lds_sync();
  }
}

// sum of the lanes' partial sums in a fixed order; every lane gets the same value
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off >= 1; off >>= 1) v = __dadd_rn(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off >= 1; off >>= 1) v = __fadd_rn(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace gwn
