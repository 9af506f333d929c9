// radix.hip -- stable LSD radix sort (8-bit digits) and exclusive scan.
//
// One pass = three kernels:
//   k_digit_hist  each 256-thread block counts the digits of its 2048-key tile
//                 in an LDS histogram, written digit-major: hist[d*nblocks+b]
//   exclusive scan of hist -> the output offset of every (digit, block)
//   k_scatter     each block re-reads its tile in order and ranks every key
//                 among the equal digits before it: inside a wavefront with
//                 8 ballots (peer mask of equal digits, popcount below the
//                 lane), across the block's 4 wavefronts with LDS counters, and
//                 across the tile's 8 rounds with running per-digit totals.
// The rank of a key is its position among equal digits in input order, so the
// sort is stable and its permutation is the one a stable CPU sort of the same
// keys gives.
#include <hip/hip_runtime.h>

#include "gwn_internal.h"

namespace gwn {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_digit_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                            uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = base + r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                         K* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n,
                                                         int shift, const uint32_t* __restrict__ offs, int nblocks) {
  __shared__ uint32_t base_off[256];   // global output offset of each digit for this tile
  __shared__ uint32_t running[256];    // keys of each digit already placed from this tile
  __shared__ uint32_t wcount[RS_WAVES][256];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  base_off[tid] = offs[(int64_t)tid * nblocks + blockIdx.x];
  running[tid] = 0;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    for (int w = 0; w < RS_WAVES; ++w) wcount[w][tid] = 0;
    __syncthreads();
    const int64_t i = base + r * RS_THREADS + tid;
    const bool valid = i < n;
    K key = valid ? kin[i] : K(0);
    uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
    // lanes of this wavefront holding the same digit
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const int bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t rank_w = (uint32_t)__popcll(peers & below);
    if (valid && rank_w == 0) wcount[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wcount[w][d];
    const uint32_t pos = base_off[d] + running[d] + before + rank_w;
    __syncthreads();
    // advance the running totals (thread t owns digit t)
    uint32_t tot = 0;
    for (int w = 0; w < RS_WAVES; ++w) tot += wcount[w][tid];
    running[tid] += tot;
    if (valid) {
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
  }
}

// ---- exclusive scan (uint32), three phases over 2048-element chunks ----
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_CHUNK = SC_THREADS * SC_ITEMS;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  // Hillis-Steele in LDS over 256 threads
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < SC_THREADS; off <<= 1) {
    uint32_t x = (t >= off) ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const uint32_t incl = sh[t];
  if (total) *total = sh[SC_THREADS - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(SC_THREADS) void k_chunk_sum(const uint32_t* __restrict__ in, int64_t n,
                                                           uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[SC_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SC_CHUNK + (int64_t)threadIdx.x * SC_ITEMS;
  uint32_t s = 0;
  for (int k = 0; k < SC_ITEMS; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  block_excl_scan(s, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_sums(uint32_t* __restrict__ sums, int64_t m,
                                                           uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[SC_THREADS];
  uint32_t carry = 0;
  for (int64_t base = 0; base < m; base += SC_THREADS) {
    const int64_t i = base + threadIdx.x;
    const uint32_t v = i < m ? sums[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan(v, sh, &tot);
    if (i < m) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SC_THREADS) void k_chunk_scan(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                            int64_t n, const uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[SC_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SC_CHUNK + (int64_t)threadIdx.x * SC_ITEMS;
  uint32_t v[SC_ITEMS];
  uint32_t s = 0;
  for (int k = 0; k < SC_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t run = sums[blockIdx.x] + block_excl_scan(s, sh, nullptr);
  for (int k = 0; k < SC_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

size_t scan_temp_bytes(int64_t n) { return sizeof(uint32_t) * ((size_t)div_up(n, SC_CHUNK) + 1); }

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev, void* tmp,
                        hipStream_t st) {
  if (n <= 0) return;
  const int64_t m = div_up(n, SC_CHUNK);
  uint32_t* sums = (uint32_t*)tmp;
  hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)m), dim3(SC_THREADS), 0, st, in, n, sums);
  check_launch("k_chunk_sum");
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_THREADS), 0, st, sums, m, total_dev);
  check_launch("k_scan_sums");
  hipLaunchKernelGGL(k_chunk_scan, dim3((unsigned)m), dim3(SC_THREADS), 0, st, in, out, n, sums);
  check_launch("k_chunk_scan");
}

// ---- radix sort driver ----
static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t radix_sort_temp_bytes(int64_t n, bool key64) {
  const int64_t nb = div_up(n > 0 ? n : 1, RS_TILE);
  const size_t hist = sizeof(uint32_t) * 256 * (size_t)nb;
  return align256((key64 ? 8 : 4) * (size_t)n) + align256(4 * (size_t)n) + 2 * align256(hist) +
         align256(scan_temp_bytes(256 * nb));
}

template <typename K>
static void radix_sort(K* keys, uint32_t* vals, int64_t n, int bits, void* tmp, size_t tmp_bytes, hipStream_t st) {
  if (n <= 1) return;
  GWN_CHECK(tmp_bytes >= radix_sort_temp_bytes(n, sizeof(K) == 8), GWN_ERR_INVALID, "radix sort temp too small");
  const int nb = (int)div_up(n, RS_TILE);
  uint8_t* w = (uint8_t*)tmp;
  K* k2 = (K*)w;
  w += align256(sizeof(K) * (size_t)n);
  uint32_t* v2 = (uint32_t*)w;
  w += align256(4 * (size_t)n);
  uint32_t* hist = (uint32_t*)w;
  w += align256(sizeof(uint32_t) * 256 * (size_t)nb);
  uint32_t* offs = (uint32_t*)w;
  w += align256(sizeof(uint32_t) * 256 * (size_t)nb);
  void* sc = w;
  K* ka = keys;
  uint32_t* va = vals;
  K* kb = k2;
  uint32_t* vb = v2;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_digit_hist<K>, dim3(nb), dim3(RS_THREADS), 0, st, ka, n, shift, hist, nb);
    check_launch("k_digit_hist");
    exclusive_scan_u32(hist, offs, 256 * (int64_t)nb, nullptr, sc, st);
    hipLaunchKernelGGL(k_scatter<K>, dim3(nb), dim3(RS_THREADS), 0, st, ka, va, kb, vb, n, shift, offs, nb);
    check_launch("k_scatter");
    std::swap(ka, kb);
    std::swap(va, vb);
    ++passes;
  }
  if (passes & 1) {
    GWN_HIP(hipMemcpyAsync(keys, ka, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, st));
    GWN_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
  }
}

void radix_sort_u32(uint32_t* keys, uint32_t* vals, int64_t n, int bits, void* tmp, size_t tmp_bytes,
                    hipStream_t st) {
  radix_sort<uint32_t>(keys, vals, n, bits, tmp, tmp_bytes, st);
}
void radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int bits, void* tmp, size_t tmp_bytes,
                    hipStream_t st) {
  radix_sort<uint64_t>(keys, vals, n, bits, tmp, tmp_bytes, st);
}

}  // namespace gwn
