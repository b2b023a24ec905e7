// k_sort.hip -- stable LSD radix sort of 30-bit keys with 32-bit payloads.
//
// Reference: scripts/bvh/kernels/radix-sort.js.  The reference runs 16
// passes of 2-bit digits, each as 4 dispatches and a host round trip
// (:386-488), with a 3-level 256-ary scan capped at 2^24 keys (:189-202).
// Its output is the unique stable-sort permutation of the keys; this file
// computes the same permutation with 4 passes of 8-bit digits:
//
//   * one upfront histogram of all four digits (fused into the Morton kernel
//     in the full build; k_histogram here for the stand-alone sort),
//   * one kernel per digit ("onesweep", Adinets & Merrill 2022): each block
//     draws a tile id from an atomic counter (CDNA does not promise block
//     issue order), ranks its 4096 keys with wave64 ballot matching into
//     per-wave LDS histograms, publishes its digit counts, and resolves its
//     global digit offsets by single-pass decoupled look-back (Merrill &
//     Garland 2016) with one thread per digit.  Keys and payloads are then
//     reordered in LDS and written out digit-run by digit-run.
//
// Look-back words are 64-bit: hi = (stamp << 2) | flag, lo = count, with
// stamp = epoch*4 + pass.  A word from an earlier pass or build carries an
// older stamp and reads as "not yet published", so the status array is never
// cleared between passes.  Loads and stores of those words are agent-scope
// atomics, so no fences are needed (count and flag travel together).
#include "device_util.h"

namespace lbvh {

constexpr int kWaves = kSortThreads / 64;
static_assert(kSortThreads == kRadix, "one thread per digit in the look-back");
constexpr uint32_t kFlagAggregate = 1u;
constexpr uint32_t kFlagInclusive = 2u;
constexpr int kLookback = 8;  // predecessor tiles probed per look-back step

__device__ __forceinline__ uint64_t lb_pack(uint32_t stamp, uint32_t flag,
                                            uint32_t count) {
  return ((uint64_t)((stamp << 2) | flag) << 32) | count;
}

// Upfront digit histograms of the stand-alone sort (bumps the epoch).
__global__ __launch_bounds__(kMapThreads) void k_histogram(
    const uint32_t *__restrict__ keys, uint32_t n, Ctl *ctl) {
  __shared__ uint32_t s_hist[kSortPasses][kRadix];
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->epoch += 1;
  for (int k = threadIdx.x; k < kSortPasses * kRadix; k += blockDim.x)
    (&s_hist[0][0])[k] = 0u;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    hist_add(s_hist, keys[i]);
  __syncthreads();
  for (int k = threadIdx.x; k < kSortPasses * kRadix; k += blockDim.x) {
    const uint32_t v = (&s_hist[0][0])[k];
    if (v) atomicAdd(&(&ctl->hist[0][0])[k], v);
  }
}

// One 8-bit digit pass.  kFirst: the payload is the input position (the
// reference's init_idx, radix-sort.js:527-534, folded into the first pass).
// With segments (batched builds), every segment is sorted on its own: tiles
// never straddle a segment, the look-back stops at the segment's first tile,
// digit bases come from the segment's histograms, and payloads are
// positions within the segment.
template <bool kFirst>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    const uint32_t *__restrict__ keys_in, const int32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, int32_t *__restrict__ vals_out,
    uint32_t n, uint32_t pass, uint32_t num_tiles, Ctl *ctl,
    uint64_t *status, int last_pass, SortSegs segs) {
  __shared__ uint32_t s_keys[kSortTile];
  __shared__ int32_t s_vals[kSortTile];
  __shared__ uint32_t s_whist[kWaves][kRadix];
  __shared__ uint32_t s_start[kRadix];  // tile-local start of each digit run
  __shared__ uint32_t s_gbase[kRadix];  // global position minus local start
  __shared__ uint32_t s_wsum[16];
  __shared__ uint32_t s_tile;
  __shared__ int s_last;

  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t shift = pass * kRadixBits;

  if (tid == 0) {
    const uint32_t t = atomicAdd(&ctl->tile_ctr, 1u);
    if (t == num_tiles - 1)
      __hip_atomic_store(&ctl->tile_ctr, 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    s_tile = t;
  }
  for (int k = tid; k < kWaves * kRadix; k += kSortThreads)
    (&s_whist[0][0])[k] = 0u;
  const uint32_t stamp = (ctl->epoch * 4u + pass) & 0x3FFFFFFFu;
  __syncthreads();
  const uint32_t tile = s_tile;
  uint32_t seg_base = 0, seg_end = n, tile0 = 0;
  const uint32_t *hist = ctl->hist[pass];
  if (segs.tile_seg) {
    const uint32_t sg = segs.tile_seg[tile];
    seg_base = segs.seg_off[sg];
    seg_end = segs.seg_off[sg + 1];
    tile0 = segs.seg_tile0[sg];
    hist = segs.seg_hist + (size_t)sg * kSortPasses * kRadix + pass * kRadix;
  }
  const uint32_t gcount = hist[tid];
  const uint32_t tile_base = seg_base + (tile - tile0) * kSortTile;

  // Warp-striped load: wave w owns keys [wbase, wbase + 64*kSortItems), item
  // j of lane l sits at wbase + 64*j + l, so (j, l) order is input order.
  const uint32_t wbase = tile_base + wave * 64 * kSortItems;
  uint32_t key[kSortItems];
  int32_t val[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t pos = wbase + 64 * j + lane;
    key[j] = 0u;
    val[j] = 0;
    if (pos < seg_end) {
      key[j] = keys_in[pos];
      val[j] = kFirst ? (int32_t)(pos - seg_base) : vals_in[pos];
    }
  }

  // Rank: for each item, the lanes holding the same digit are found with 8
  // ballots; the lowest of them adds the group size to the wave's digit
  // counter.  rank = counter before the add + position inside the group.
  uint32_t rank[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const bool valid = (wbase + 64 * j + lane) < seg_end;
    const uint32_t d = (key[j] >> shift) & (kRadix - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t vote = __ballot(bit);
      peers &= bit ? vote : ~vote;
    }
    rank[j] = 0u;
    if (valid) {
      const uint32_t below = popc_below(peers);
      const uint32_t prev = s_whist[wave][d];
      if (below == 0) s_whist[wave][d] = prev + (uint32_t)__popcll(peers);
      rank[j] = prev + below;
    }
  }
  __syncthreads();

  // Thread d: this tile's count of digit d, per-wave exclusive offsets.
  const uint32_t d = tid;
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const uint32_t c = s_whist[w][d];
    s_whist[w][d] = total;
    total += c;
  }
  uint64_t *my_word = status + (size_t)tile * kRadix + d;
  __hip_atomic_store(my_word,
                     lb_pack(stamp, tile == tile0 ? kFlagInclusive : kFlagAggregate,
                             total),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

  const uint32_t start = block_exclusive_scan(total, s_wsum);
  const uint32_t gdigit = block_exclusive_scan(gcount, s_wsum);
  s_start[d] = start;

  // Decoupled look-back for digit d, kLookback predecessor words per probe:
  // the loads of one probe are independent, so a walk over tiles that have
  // only published aggregates advances kLookback tiles per memory round
  // trip.  Every earlier tile drew its id before this one and publishes its
  // aggregate without waiting, so the walk ends; the spin cap only bounds the
  // loop should that invariant ever break (the result is then wrong and
  // ctl->fault says so, but the kernel still exits).
  uint32_t excl = 0;
  if (tile > tile0) {
    int64_t k = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (k >= (int64_t)tile0) {
      uint64_t w[kLookback];
#pragma unroll
      for (int j = 0; j < kLookback; ++j)
        w[j] = (k - j >= (int64_t)tile0)
                   ? __hip_atomic_load(status + (size_t)(k - j) * kRadix + d,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : 0ull;
      int used = 0;
      bool done = false;
#pragma unroll
      for (int j = 0; j < kLookback; ++j) {
        if (done || used < j) continue;      // stopped earlier in this probe
        if (k - j < (int64_t)tile0) { done = true; continue; }
        const uint32_t hi = (uint32_t)(w[j] >> 32);
        if ((hi >> 2) != stamp) continue;  // not published yet
        excl += (uint32_t)w[j];
        used = j + 1;
        if ((hi & 3u) == kFlagInclusive) done = true;
      }
      if (done) break;
      k -= used;
      if (used == 0) {
        if (++spins > kSpinCap) {
          atomicOr(&ctl->fault, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(my_word, lb_pack(stamp, kFlagInclusive, excl + total),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  s_gbase[d] = seg_base + gdigit + excl - start;
  __syncthreads();

  // Reorder the tile in LDS by (digit, input order).
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    if ((wbase + 64 * j + lane) < seg_end) {
      const uint32_t dj = (key[j] >> shift) & (kRadix - 1);
      const uint32_t lp = s_start[dj] + s_whist[wave][dj] + rank[j];
      s_keys[lp] = key[j];
      s_vals[lp] = val[j];
    }
  }
  __syncthreads();

  // Write out: consecutive threads take consecutive tile positions, so each
  // digit run goes to consecutive global addresses.
  const uint32_t tile_n = (seg_end - tile_base < (uint32_t)kSortTile)
                              ? seg_end - tile_base : (uint32_t)kSortTile;
  for (uint32_t p = tid; p < tile_n; p += kSortThreads) {
    const uint32_t k = s_keys[p];
    const uint32_t g = s_gbase[(k >> shift) & (kRadix - 1)] + p;
    if (g >= seg_base && g < seg_end) {
      keys_out[g] = k;
      vals_out[g] = s_vals[p];
    }
  }

  if (last_pass && !segs.tile_seg) {
    // The last block to finish clears the digit histograms for the next
    // operation: every block read its counts before arriving here.  (Batched
    // builds clear their per-segment histograms up front.)
    if (tid == 0) {
      const uint32_t done = __hip_atomic_fetch_add(
          &ctl->done_ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_last = (done == num_tiles - 1);
    }
    __syncthreads();
    if (s_last) {
      for (int k = tid; k < kSortPasses * kRadix; k += kSortThreads)
        (&ctl->hist[0][0])[k] = 0u;
      if (tid == 0)
        __hip_atomic_store(&ctl->done_ctr, 0u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

void launch_histogram(const uint32_t *keys, uint32_t n, Ctl *ctl,
                      hipStream_t s) {
  uint32_t blocks = (n + kMapThreads - 1) / kMapThreads;
  if (blocks > (uint32_t)kMortonBlocksMax) blocks = kMortonBlocksMax;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_histogram, dim3(blocks), dim3(kMapThreads), 0, s, keys,
                     n, ctl);
}

void launch_onesweep_pass(int pass, const uint32_t *keys_in,
                          const int32_t *vals_in, uint32_t *keys_out,
                          int32_t *vals_out, uint32_t n, Ctl *ctl,
                          uint64_t *status, hipStream_t s,
                          const SortSegs *segs, uint32_t seg_tiles) {
  const SortSegs one{};
  const SortSegs &sg = segs ? *segs : one;
  const uint32_t tiles = segs ? seg_tiles : sort_tiles(n);
  const int last = (pass == kSortPasses - 1) ? 1 : 0;
  if (pass == 0)
    hipLaunchKernelGGL(k_onesweep<true>, dim3(tiles), dim3(kSortThreads), 0, s,
                       keys_in, (const int32_t *)nullptr, keys_out, vals_out,
                       n, 0u, tiles, ctl, status, last, sg);
  else
    hipLaunchKernelGGL(k_onesweep<false>, dim3(tiles), dim3(kSortThreads), 0,
                       s, keys_in, vals_in, keys_out, vals_out, n,
                       (uint32_t)pass, tiles, ctl, status, last, sg);
}

}  // namespace lbvh
