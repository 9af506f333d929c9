// radix.hip -- stable LSD radix sort (8-bit digits) and exclusive scan.
//
// One pass = three kernels, none waiting on another block:
//   k_rs_tile_hist    each 256-thread block counts the digits of its 2048-key
//                     tile in LDS and writes them digit-major:
//                     counts[d * ntiles + tile]
//   k_rs_digit_scan   one block per digit: exclusive scan of that digit's
//                     counts over the tiles (where the tile's keys of digit d
//                     go among all keys of digit d), and the digit's total
//   k_rs_scatter      each block scans the 256 digit totals (where each
//                     digit's run starts), ranks every key of its tile among
//                     the equal digits before it (8 ballots give the peer mask
//                     inside a wavefront; per-wavefront LDS counters carry it
//                     across the wavefront's 8 rounds, a prefix over the 4
//                     wavefronts across the tile), stages the tile in LDS in
//                     output order and writes it out in digit runs (coalesced).
// The rank of a key is its position among equal digits in input order, so the
// sort is stable and its permutation is the one a stable CPU sort of the same
// keys gives.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gwn_internal.h"

namespace gwn {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_tile_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                              uint32_t* __restrict__ counts, int64_t nt) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = base + r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  counts[(int64_t)threadIdx.x * nt + blockIdx.x] = h[threadIdx.x];
}

// inclusive scan over a 256-thread block: shuffles inside each wavefront,
// then the wavefront totals (one barrier); `wsum` holds RS_WAVES words
__device__ __forceinline__ uint32_t block_incl_scan256(uint32_t v, uint32_t* wsum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t x = __shfl_up(v, off, 64);
    if (lane >= off) v += x;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  for (int w = 0; w < wave; ++w) v += wsum[w];
  __syncthreads();
  return v;
}

// block d: exclusive scan of counts[d * nt .. d * nt + nt) in place; totals[d]
__global__ __launch_bounds__(256) void k_rs_digit_scan(uint32_t* __restrict__ counts, int64_t nt,
                                                       uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[RS_WAVES];
  __shared__ uint32_t last;
  const int t = threadIdx.x;
  uint32_t* c = counts + (int64_t)blockIdx.x * nt;
  uint32_t carry = 0;
  for (int64_t b = 0; b < nt; b += 256) {
    const int64_t i = b + t;
    const uint32_t v = i < nt ? c[i] : 0u;
    const uint32_t incl = block_incl_scan256(v, wsum);
    if (i < nt) c[i] = carry + incl - v;
    if (t == 255) last = incl;
    __syncthreads();
    carry += last;
    __syncthreads();
  }
  if (t == 0) totals[blockIdx.x] = carry;
}

// Keys of a tile in wavefront-major order: wavefront w takes keys
// [w * 512, (w + 1) * 512) of the tile in 8 rounds of 64 consecutive keys,
// so (wavefront, round, lane) is input order and a key's rank among equal
// digits needs no barrier inside the tile's ranking: each wavefront keeps its
// own digit counters in LDS.
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                            K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                            int64_t n, int shift, const uint32_t* __restrict__ counts,
                                                            const uint32_t* __restrict__ totals, int64_t nt) {
  __shared__ uint32_t wh[RS_WAVES][256];  // per wavefront: keys of each digit so far; then their prefix over wavefronts
  __shared__ uint32_t wsum[RS_WAVES];
  __shared__ uint32_t loff[256];          // where each digit's run starts in the tile
  __shared__ uint32_t gbase[256];         // where it starts in the output
  __shared__ K s_keys[RS_TILE];
  __shared__ uint32_t s_vals[RS_TILE];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * RS_TILE;
  for (int w = 0; w < RS_WAVES; ++w) wh[w][tid] = 0;
  // where each digit's run starts in the output: exclusive scan of the totals
  const uint32_t tot_d = totals[tid];
  const uint32_t gpre = block_incl_scan256(tot_d, wsum) - tot_d;  // (its barriers also publish wh's zeros)
  gbase[tid] = gpre + counts[(int64_t)tid * nt + tile];
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t wbase = base + (int64_t)wave * (64 * RS_ITEMS);
  K key[RS_ITEMS];
  uint32_t val[RS_ITEMS], lrank[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = wbase + r * 64 + lane;
    const bool valid = i < n;
    key[r] = valid ? kin[i] : K(0);
    val[r] = valid ? vin[i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const bool valid = wbase + r * 64 + lane < n;
    const uint32_t d = (uint32_t)(key[r] >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const int bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank_w = (uint32_t)__popcll(peers & below);
    // every lane reads its digit's count before the group's first lane adds
    // the group to it (one wavefront: the LDS read completes first)
    lrank[r] = wh[wave][d] + rank_w;
    if (valid && rank_w == 0) wh[wave][d] += (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // thread tid owns digit tid: tile count and the prefix over wavefronts
  uint32_t mine = 0;
  for (int w = 0; w < RS_WAVES; ++w) {
    const uint32_t c = wh[w][tid];
    wh[w][tid] = mine;
    mine += c;
  }
  loff[tid] = block_incl_scan256(mine, wsum) - mine;
  __syncthreads();
  // stage in tile output order
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    if (wbase + r * 64 + lane < n) {
      const uint32_t d = (uint32_t)(key[r] >> shift) & 0xFFu;
      const uint32_t lp = loff[d] + wh[wave][d] + lrank[r];
      s_keys[lp] = key[r];
      s_vals[lp] = val[r];
    }
  }
  __syncthreads();
  // write out: consecutive threads take consecutive keys of a digit run
  const int64_t cnt = (n - base) < RS_TILE ? (n - base) : RS_TILE;
  for (int j = tid; j < cnt; j += RS_THREADS) {
    const K k = s_keys[j];
    const uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
    const uint32_t g = gbase[d] + (uint32_t)(j - (int)loff[d]);
    kout[g] = k;
    vout[g] = s_vals[j];
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

// workspace: the ping-pong keys and values, the digit-major tile counts and
// the 256 digit totals
size_t radix_sort_temp_bytes(int64_t n, bool key64) {
  const int64_t nt = div_up(n > 0 ? n : 1, RS_TILE);
  return align256((key64 ? 8 : 4) * (size_t)n) + align256(4 * (size_t)n) + align256(4 * 256 * (size_t)nt) +
         align256(4 * 256);
}

template <typename K>
static void radix_sort(K* keys, uint32_t* vals, int64_t n, int bits, void* tmp, size_t tmp_bytes, hipStream_t st) {
  if (n <= 1) return;
  GWN_CHECK(tmp_bytes >= radix_sort_temp_bytes(n, sizeof(K) == 8), GWN_ERR_INVALID, "radix sort temp too small");
  GWN_CHECK(n < ((int64_t)1 << 32), GWN_ERR_INVALID, "radix sort: more than 2^32 - 1 keys");
  const int64_t nt = div_up(n, RS_TILE);
  uint8_t* w = (uint8_t*)tmp;
  K* k2 = (K*)w;
  w += align256(sizeof(K) * (size_t)n);
  uint32_t* v2 = (uint32_t*)w;
  w += align256(4 * (size_t)n);
  uint32_t* counts = (uint32_t*)w;
  w += align256(4 * 256 * (size_t)nt);
  uint32_t* totals = (uint32_t*)w;
  K* ka = keys;
  uint32_t* va = vals;
  K* kb = k2;
  uint32_t* vb = v2;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_tile_hist<K>, dim3((unsigned)nt), dim3(RS_THREADS), 0, st, ka, n, shift, counts, nt);
    check_launch("k_rs_tile_hist");
    hipLaunchKernelGGL(k_rs_digit_scan, dim3(256), dim3(256), 0, st, counts, nt, totals);
    check_launch("k_rs_digit_scan");
    hipLaunchKernelGGL(k_rs_scatter<K>, dim3((unsigned)nt), dim3(RS_THREADS), 0, st, ka, va, kb, vb, n, shift, counts,
                       totals, nt);
    check_launch("k_rs_scatter");
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
