// k_sort.hip -- stable LSD radix sort of 30-bit keys with 32-bit payloads.
//
// Reference: scripts/bvh/kernels/radix-sort.js.  The reference runs 16
// passes of 2-bit digits, each as 4 dispatches and a host round trip
// (:386-488), with a 3-level 256-ary scan capped at 2^24 keys (:189-202).
// Its output is the unique stable-sort permutation of the keys; this file
// computes the same permutation with 4 passes of 8-bit digits, or 3 passes
// of 10-bit digits (the digit width is a tunable: wider digits take fewer
// passes and launches, narrower ones less LDS per block and fewer count
// words per tile).
//
// Each pass is reduce-then-scan over tiles of kSortTile keys (kSmallSortTile
// for a single mesh of at most kSmallSortMax keys):
//   count    per-tile digit counts (pass 0's are written by the Morton
//            kernel as it produces the keys), and their column sums;
//   scan     one block per digit scans that digit's column of tile counts
//            -> each (digit, tile) run's position among the digit's keys,
//            and the column total (the global histogram);
//            the scatter adds the digit's base (an exclusive scan of the
//            totals, which every tile computes while its keys load);
//   scatter  each tile ranks its keys with wave64 ballot matching into
//            per-wave LDS histograms (rank order = input order: stable),
//            reorders them in LDS and writes each digit run contiguously.
// No block ever waits for another block.  A decoupled look-back chains every
// tile to its predecessors, and on MI355X each link of that chain is a
// coherent cross-XCD load; with thousands of tiles the chain, not HBM, set
// the pass time.  Kernel boundaries are the only global synchronisation.
//
// A pass's global digit histogram (the digit bases of a single mesh) is the
// column sums of its per-tile counts.  A batched sort's meshes are segments
// of whole tiles, each sorted on its own: one block per mesh turns its
// tiles' counts into their offsets (k_seg_scan).  No device-scope atomics
// anywhere.  A single mesh's counts and offsets are digit-major, entry
// (d, t) at d * tiles + t, so the column scan reads and writes columns
// contiguously; a batched sort's are tile-major, (d, t) at t * R + d, so a
// mesh's entries are one contiguous run however few tiles it has (digit-
// major, a mesh of a tile or two touched a cache line per digit for a few
// bytes each).  Tile ids are XCD-aware
// (xcd_tile): consecutive tiles run on the same XCD, so the cache lines that
// neighbouring tiles share -- count columns and the boundaries of their digit
// runs in the output -- are merged in one L2.
#include "device_util.h"

namespace lbvh {

constexpr int kWaves = kSortThreads / 64;
// A column scan's block: 1,024 threads, or 256 (64) for at most kScanSmall
// (kScanSmall / 4) tiles, a column that fits one chunk of the smaller block.
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;  // column entries per thread per scan chunk
constexpr uint32_t kScanSmall = 256 * kScanPer;
// Single-mesh sorts of at most this many tiles (64K keys) launch no scan:
// each scatter tile sums the count rows of its digits itself.
constexpr uint32_t kSmallScanTiles = 16;

// Digits of width kBits: kR values, kDpt of them per thread of a block of
// kSortThreads (thread t owns digits [t * kDpt, (t + 1) * kDpt)).
template <int kBits>
struct Digits {
  static constexpr int kR = 1 << kBits;
  static constexpr int kDpt = kR / kSortThreads;
  static_assert(kDpt >= 1 && kR % kSortThreads == 0, "whole digits per thread");
};

// Tile loads of a large sort are non-temporal: the keys and payloads are
// read once per pass, from a buffer far larger than the L2s, and would
// push out lines that are still read.  A sort of small tiles (at most
// kSmallSortMax keys) keeps plain loads: its keys are still in cache from
// the pass before, and non-temporal loads made a 64K build 2 us slower.
template <int kItems>
constexpr bool kStreamNt = kItems == kSortItems;

// A digit of at most kFewBits significant bits (the last one of Morton
// keys) takes at most kFewVals values, and few of them per tile: counters
// that the lanes of a wave share would serialise its atomics.  Each lane
// then counts into a row of its own (rows of kFewVals + 1 words, so lanes
// adding to one digit hit distinct banks), and the block sums the rows.
constexpr int kFewBits = 6;
constexpr int kFewVals = 1 << kFewBits;
constexpr int kFewRow = kFewVals + 1;

// Adds a valid digit to the block's counters (kFew: to the calling lane's
// row; the mask only keeps the row in bounds, the digit has no higher bits).
template <int kBits, bool kFew>
__device__ __forceinline__ void count_digit(uint32_t *s_c, uint32_t d,
                                            bool valid) {
  if (kFew) {
    if (valid)
      atomicAdd(&s_c[(threadIdx.x & 63) * kFewRow + (d & (kFewVals - 1))], 1u);
  } else if (valid) {
    digit_add(s_c, d);
  }
}

// Per-tile digit counts of one pass -> cnt (digit-major, or tile-major for a
// batched sort).  kFew: a digit of at most kFewBits significant bits.
template <int kBits, bool kFew, int kItems>
__global__ __launch_bounds__(kSortThreads) void k_count(
    const uint32_t *__restrict__ keys, uint32_t n, uint32_t shift,
    uint32_t *__restrict__ cnt, SortSegs segs) {
  using D = Digits<kBits>;
  static_assert(!kFew || (kFewVals <= D::kR && kSortThreads % kFewVals == 0),
                "a few-bit digit's values are digits; whole row groups");
  constexpr int kWords = kFew ? 64 * kFewRow : D::kR;
  __shared__ uint32_t s_c[kWords];
  __shared__ uint32_t s_few[kFewVals];  // the rows' sums (kFew)
  const uint32_t tiles = gridDim.x;
  const uint32_t tile = xcd_tile(blockIdx.x, tiles);
  for (int w = threadIdx.x; w < kWords; w += kSortThreads) s_c[w] = 0u;
  if (kFew && threadIdx.x < kFewVals) s_few[threadIdx.x] = 0u;
  __syncthreads();
  const TileRange r = sort_tile_range(tile, n, segs, kSortThreads * kItems);
  // All of the thread's keys are loaded before any is counted, so a wave
  // has its whole share of the tile in flight; a whole 16-byte-aligned tile
  // with 16-byte loads (counting does not depend on which thread holds
  // which key).
  uint32_t k[kItems];
  const bool vec =
      r.end - r.begin == (uint32_t)(kSortThreads * kItems) &&
      (reinterpret_cast<uintptr_t>(keys + r.begin) & 15u) == 0;
  if (vec) {
    const uint4 *k4 = reinterpret_cast<const uint4 *>(keys + r.begin);
#pragma unroll
    for (int q = 0; q < kItems / 4; ++q) {
      const uint4 *src = k4 + threadIdx.x + q * kSortThreads;
      const uint4 v = kStreamNt<kItems> ? load_nt(src) : *src;
      k[4 * q] = v.x;
      k[4 * q + 1] = v.y;
      k[4 * q + 2] = v.z;
      k[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      count_digit<kBits, kFew>(s_c, (k[j] >> shift) & (D::kR - 1), true);
  } else {
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const uint32_t i = r.begin + threadIdx.x + j * kSortThreads;
      k[j] = i < r.end ? keys[i] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      count_digit<kBits, kFew>(s_c, (k[j] >> shift) & (D::kR - 1),
                               r.begin + threadIdx.x + j * kSortThreads < r.end);
  }
  __syncthreads();
  if (kFew) {
    // thread t sums rows [g * kRows, (g + 1) * kRows) of value t % kFewVals,
    // g = t / kFewVals
    constexpr int kGroups = kSortThreads / kFewVals;
    constexpr int kRows = 64 / kGroups;
    const int v = threadIdx.x % kFewVals, g = threadIdx.x / kFewVals;
    uint32_t sum = 0;
#pragma unroll
    for (int l = 0; l < kRows; ++l) sum += s_c[(g * kRows + l) * kFewRow + v];
    atomicAdd(&s_few[v], sum);
    __syncthreads();
  }
  const bool tile_major = segs.tile_seg != nullptr;  // batched
  for (int d = threadIdx.x; d < D::kR; d += kSortThreads) {
    const uint32_t c = kFew ? (d < kFewVals ? s_few[d] : 0u) : s_c[d];
    cnt[tile_major ? (size_t)tile * D::kR + d : (size_t)d * tiles + tile] = c;
  }
}

// Exclusive prefix of h over all kR digits at each of the calling thread's
// kDpt digits (digit t * kDpt + i -> out[i]); sum = the thread's total.
template <int kBits>
__device__ __forceinline__ void digit_prefix(const uint32_t (&h)[Digits<kBits>::kDpt],
                                             uint32_t (&out)[Digits<kBits>::kDpt],
                                             uint32_t *s_wsum) {
  using D = Digits<kBits>;
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < D::kDpt; ++i) sum += h[i];
  uint32_t run = block_exclusive_scan(sum, s_wsum);
#pragma unroll
  for (int i = 0; i < D::kDpt; ++i) {
    out[i] = run;
    run += h[i];
  }
}

// Column scan of a single mesh: block d turns column d of the tile counts
// into output positions, off[d][t] = P(t) = sum of cnt[d][t'] over t' < t,
// and hist[d] = the column total; the scatter adds base(d) = keys with a
// smaller digit (an exclusive scan of hist over the digits, which needs
// every column).  Chunks of kThreads * kScanPer column entries go
// through LDS, so global loads and stores are coalesced while each thread
// scans kScanPer consecutive entries (one pad word per kScanPer keeps the
// per-thread reads conflict-free).  Four entries per thread and 1,024
// threads took a 16M pass's scan from 7.8 to about 4 us against sixteen
// and 256: the block's serial part is a quarter as long.
template <int kBits, int kThreads>
__global__ __launch_bounds__(kThreads) void k_scan_columns(
    const uint32_t *__restrict__ cnt, uint32_t *__restrict__ off,
    uint32_t tiles, uint32_t *__restrict__ hist) {
  constexpr int kChunk = kThreads * kScanPer;
  __shared__ uint32_t s_col[kChunk + kChunk / kScanPer];
  __shared__ uint32_t s_wsum[16];
  __shared__ uint32_t s_carry;
  const uint32_t d = blockIdx.x;
  uint32_t carry = 0;
  const uint32_t *col = cnt + (size_t)d * tiles;
  uint32_t *out = off + (size_t)d * tiles;
  const auto pad = [](uint32_t e) { return e + e / kScanPer; };
  for (uint32_t c0 = 0; c0 < tiles; c0 += kChunk) {
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
      const uint32_t e = j * kThreads + threadIdx.x;
      s_col[pad(e)] = (c0 + e < tiles) ? col[c0 + e] : 0u;
    }
    __syncthreads();
    uint32_t v[kScanPer];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
      v[j] = s_col[pad(threadIdx.x * kScanPer + j)];
      sum += v[j];
    }
    const uint32_t ex = block_exclusive_scan(sum, s_wsum);
    if (threadIdx.x == kThreads - 1) s_carry = ex + sum;
    uint32_t run = carry + ex;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
      s_col[pad(threadIdx.x * kScanPer + j)] = run;
      run += v[j];
    }
    __syncthreads();  // s_carry and the staged offsets are visible
    carry += s_carry;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
      const uint32_t e = j * kThreads + threadIdx.x;
      if (c0 + e < tiles) out[c0 + e] = s_col[pad(e)];
    }
    __syncthreads();  // s_col and s_carry are rewritten by the next chunk
  }
  if (threadIdx.x == 0) hist[d] = carry;
}

// Offsets of a batched pass, one block per segment (mesh).  A segment's
// tiles are consecutive and hold only its keys, so the run of digit d in
// its tile t starts at seg_off + (the segment's keys with a smaller digit)
// + (digit-d keys in the segment's earlier tiles).  The counts are
// tile-major, so the segment's are one contiguous run of tiles * R words:
// thread t owns digits [t * kDpt, (t + 1) * kDpt) and reads them tile after
// tile (each tile's R words in one coalesced sweep of the block), kUnroll
// tiles in flight; it sums them over the segment, the block scans the sums
// into bases, and the thread writes its running positions.  An empty
// segment has no tiles.
template <int kBits>
__global__ __launch_bounds__(kSortThreads) void k_seg_scan(
    const uint32_t *__restrict__ cnt, uint32_t *__restrict__ off,
    uint32_t tiles, SortSegs segs, uint32_t num_segs) {
  using D = Digits<kBits>;
  constexpr int kUnroll = 8;
  __shared__ uint32_t s_wsum[16];
  const uint32_t sg = blockIdx.x;
  const uint32_t t0 = segs.seg_tile0[sg];
  const uint32_t t1 = (sg + 1 < num_segs) ? segs.seg_tile0[sg + 1] : tiles;
  const uint32_t d0 = threadIdx.x * D::kDpt;
  uint32_t total[D::kDpt];
#pragma unroll
  for (int i = 0; i < D::kDpt; ++i) total[i] = 0u;
  for (uint32_t t = t0; t < t1; t += kUnroll) {
    uint32_t c[kUnroll][D::kDpt];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int i = 0; i < D::kDpt; ++i)
        c[u][i] = t + u < t1 ? cnt[(size_t)(t + u) * D::kR + d0 + i] : 0u;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int i = 0; i < D::kDpt; ++i) total[i] += c[u][i];
  }
  uint32_t run[D::kDpt];
  digit_prefix<kBits>(total, run, s_wsum);
  const uint32_t first = segs.seg_off[sg];
#pragma unroll
  for (int i = 0; i < D::kDpt; ++i) run[i] += first;
  for (uint32_t t = t0; t < t1; t += kUnroll) {
    uint32_t c[kUnroll][D::kDpt];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int i = 0; i < D::kDpt; ++i)
        c[u][i] = t + u < t1 ? cnt[(size_t)(t + u) * D::kR + d0 + i] : 0u;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (t + u >= t1) break;
#pragma unroll
      for (int i = 0; i < D::kDpt; ++i) {
        off[(size_t)(t + u) * D::kR + d0 + i] = run[i];
        run[i] += c[u][i];
      }
    }
  }
}

// digit_prefix of two counts at once, packed (a << 32) | b: the b sums
// (digit totals, at most 2^32 - 1 keys in all) never carry into a.
template <int kBits>
__device__ __forceinline__ void digit_prefix2(
    const uint64_t (&h)[Digits<kBits>::kDpt],
    uint64_t (&out)[Digits<kBits>::kDpt], uint64_t *s_wsum) {
  using D = Digits<kBits>;
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < D::kDpt; ++i) sum += h[i];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_wsum[wave] = x;
  __syncthreads();
  uint64_t run = x - sum;
  for (uint32_t w = 0; w < wave; ++w) run += s_wsum[w];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < D::kDpt; ++i) {
    out[i] = run;
    run += h[i];
  }
}

// One digit pass.  kFirst: the payload is the position in the segment (the
// reference's init_idx, radix-sort.js:527-534, folded into the first pass).
// hist (or null when off holds whole output positions): the pass's digit
// totals; off then holds positions among the digit's keys and each tile
// adds the digit bases, an exclusive scan of hist, computed in the same
// block scan as the tile's own run starts.  small (a single mesh of at most
// kSmallScanTiles tiles): off holds the per-tile counts themselves, and each
// tile sums its digits' rows (from L2) into its column prefix and the digit
// totals, so a small sort launches no scan kernel.
template <int kBits, bool kFirst, int kItems>
__global__ __launch_bounds__(kSortThreads) void k_scatter(
    const uint32_t *__restrict__ keys_in, const int32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, int32_t *__restrict__ vals_out,
    uint32_t n, uint32_t shift, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ hist, int small, SortSegs segs) {
  using D = Digits<kBits>;
  constexpr int kTile = kSortThreads * kItems;
  static_assert(kItems % 4 == 0, "16-byte tile loads");
  // Wide digits keep two 16-bit per-wave counters per word (a wave counts
  // at most 64 * kItems = 1024 keys of a digit, so no carry crosses into
  // the other half): 8 KB less LDS for 10-bit digits, a third block per CU.
  constexpr bool kPack = kBits > kRadixBits;
  constexpr uint32_t kHw = kPack ? D::kR / 2 : D::kR;  // counter words per wave
  static_assert(!kPack || (64 * kItems < 65536 && D::kDpt % 2 == 0),
                "16-bit counters, a thread's digits in whole words");
  __shared__ uint32_t s_keys[kTile];
  __shared__ int32_t s_vals[kTile];
  __shared__ uint32_t s_whist[kWaves][kHw];
  __shared__ uint32_t s_spare[kWaves][64];  // targets of the no-op adds
  __shared__ uint32_t s_start[D::kR];  // tile-local start of each digit run
  // the digit totals (hist), then output position minus local start
  __shared__ uint32_t s_gbase[D::kR];
  __shared__ uint32_t s_wsum[16];
  __shared__ uint64_t s_wsum2[16];

  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t tiles = gridDim.x;
  const uint32_t tile = xcd_tile(blockIdx.x, tiles);
  for (int k = tid; k < kWaves * (int)kHw; k += kSortThreads)
    (&s_whist[0][0])[k] = 0u;
  // The digit totals go straight from global memory into s_gbase (no
  // registers held meanwhile): wave w's q-th load fills 64 consecutive
  // words at (q * kWaves + w) * 64.
  if (hist) {
#pragma unroll
    for (int q = 0; q < D::kDpt; ++q) {
      const uint32_t e = (q * kWaves + wave) * 64;
      __builtin_amdgcn_global_load_lds(hist + e + lane, s_gbase + e, 4, 0, 0);
    }
  }
  uint32_t goff[D::kDpt];
  if (small) {
    // column prefix at this tile and the digit's total, from its count row
#pragma unroll
    for (int i = 0; i < D::kDpt; ++i) {
      const uint32_t d = tid * D::kDpt + i;
      const uint32_t *row = off + (size_t)d * tiles;
      uint32_t run = 0, pre = 0;
#pragma unroll
      for (uint32_t t = 0; t < kSmallScanTiles; ++t) {
        if (t < tiles) {
          if (t == tile) pre = run;
          run += row[t];
        }
      }
      goff[i] = pre;
      s_gbase[d] = run;  // read back by this thread only
    }
  } else if (segs.tile_seg) {  // batched: tile-major
#pragma unroll
    for (int i = 0; i < D::kDpt; ++i)
      goff[i] = off[(size_t)tile * D::kR + tid * D::kDpt + i];
  } else {
#pragma unroll
    for (int i = 0; i < D::kDpt; ++i)
      goff[i] = off[(size_t)(tid * D::kDpt + i) * tiles + tile];
  }
  const TileRange r = sort_tile_range(tile, n, segs, kTile);
  // No barrier here: one waits for every outstanding global load, and the
  // tile's own loads below should be in flight with the ones above.  The
  // zeroed counters and s_gbase are first used after the next barrier.

  // Warp-striped keys: wave w owns keys [wbase, wbase + 64*kItems), item
  // j of lane l sits at wbase + 64*j + l, so (j, l) order is input order.
  // A whole 16-byte-aligned tile is loaded with 16-byte loads into the LDS
  // reorder buffer (free until the reorder) and its keys read back striped:
  // 4 loads per thread instead of 16 (keys) + 16 (payloads).  Payloads stay
  // in s_vals (at the striped position) until the ranking is done, so they
  // hold no registers while it runs.
  const uint32_t wbase = r.begin + wave * 64 * kItems;
  uint32_t key[kItems];
  const bool vec =
      r.end - r.begin == (uint32_t)kTile &&
      ((reinterpret_cast<uintptr_t>(keys_in + r.begin) |
        (kFirst ? 0u : reinterpret_cast<uintptr_t>(vals_in + r.begin))) &
       15u) == 0;
  if (vec) {
    const uint4 *k4 = reinterpret_cast<const uint4 *>(keys_in + r.begin);
    const int4 *v4 = reinterpret_cast<const int4 *>(vals_in + r.begin);
    uint4 kq[kItems / 4];
    int4 vq[kItems / 4];
#pragma unroll
    for (int q = 0; q < kItems / 4; ++q) {
      kq[q] = kStreamNt<kItems> ? load_nt(k4 + tid + q * kSortThreads)
                                : k4[tid + q * kSortThreads];
      if (!kFirst)
        vq[q] = kStreamNt<kItems> ? load_nt(v4 + tid + q * kSortThreads)
                                  : v4[tid + q * kSortThreads];
    }
#pragma unroll
    for (int q = 0; q < kItems / 4; ++q) {
      reinterpret_cast<uint4 *>(s_keys)[tid + q * kSortThreads] = kq[q];
      if (!kFirst)
        reinterpret_cast<int4 *>(s_vals)[tid + q * kSortThreads] = vq[q];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      key[j] = s_keys[wave * 64 * kItems + 64 * j + lane];
  } else {
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const uint32_t pos = wbase + 64 * j + lane;
      key[j] = 0u;
      if (pos < r.end) {
        key[j] = keys_in[pos];
        if (!kFirst)
          s_vals[wave * 64 * kItems + 64 * j + lane] = vals_in[pos];
      }
    }
    __syncthreads();  // the zeroed counters
  }

  // Rank: for each item, the lanes holding the same digit are found with
  // kBits ballots; the lowest of them adds the group size to the wave's
  // digit counter with an LDS atomic that returns the count before the add,
  // and the group reads it from that lane.  rank = that count + position
  // inside the group.  A wave's LDS operations complete in issue order, so
  // all the items' atomics are issued first and their results read
  // afterwards (rank[j] holds the position in bits 0-7 and the group's first
  // lane in bits 8-13 meanwhile).
  uint32_t rank[kItems], prev[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const bool valid = (wbase + 64 * j + lane) < r.end;
    const uint32_t d = (key[j] >> shift) & (D::kR - 1);
    const uint64_t peers = match_digit<kBits>(d, valid);
    const uint32_t below = popc_below(peers);
    // Every lane issues the atomic (no branch, so nothing waits for its
    // result here): a group's first lane adds the group size to the digit
    // counter, the others add 0 to a word of their own.
    const bool lead = valid && below == 0;
    uint32_t *slot =
        lead ? &s_whist[wave][kPack ? d >> 1 : d] : &s_spare[wave][lane];
    prev[j] = atomicAdd(slot, lead ? (uint32_t)__popcll(peers)
                                         << (kPack ? (d & 1) << 4 : 0)
                                   : 0u);
    // a valid lane's group holds itself, so peers != 0
    const uint32_t leader =
        valid ? (uint32_t)__ffsll((unsigned long long)peers) - 1 : lane;
    rank[j] = below | (leader << 8);
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    uint32_t p = __shfl(prev[j], (int)(rank[j] >> 8), 64);
    if (kPack) p = (p >> ((((key[j] >> shift) & 1u)) << 4)) & 0xffffu;
    rank[j] = (wbase + 64 * j + lane) < r.end ? p + (rank[j] & 0xffu) : 0u;
  }
  __syncthreads();

  // This tile's count of each of the thread's digits and per-wave exclusive
  // offsets, then the tile-local start of each digit run.
  uint32_t total[D::kDpt], start[D::kDpt];
  if (kPack) {
    // the thread's digits pair up in whole words
#pragma unroll
    for (int i = 0; i < D::kDpt; i += 2) {
      const uint32_t q = (tid * D::kDpt + i) >> 1;
      total[i] = total[i + 1] = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = s_whist[w][q];
        s_whist[w][q] = total[i] | (total[i + 1] << 16);
        total[i] += c & 0xffffu;
        total[i + 1] += c >> 16;
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < D::kDpt; ++i) {
      const uint32_t d = tid * D::kDpt + i;
      total[i] = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = s_whist[w][d];
        s_whist[w][d] = total[i];
        total[i] += c;
      }
    }
  }
  if (hist || small) {
    // starts and digit bases in one scan; each thread reads and rewrites
    // only its own digits' s_gbase words
    uint64_t two[D::kDpt], pre[D::kDpt];
#pragma unroll
    for (int i = 0; i < D::kDpt; ++i)
      two[i] = ((uint64_t)total[i] << 32) | s_gbase[tid * D::kDpt + i];
    digit_prefix2<kBits>(two, pre, s_wsum2);
#pragma unroll
    for (int i = 0; i < D::kDpt; ++i) {
      start[i] = (uint32_t)(pre[i] >> 32);
      goff[i] += (uint32_t)pre[i];
    }
  } else {
    digit_prefix<kBits>(total, start, s_wsum);
  }
#pragma unroll
  for (int i = 0; i < D::kDpt; ++i) {
    s_start[tid * D::kDpt + i] = start[i];
    s_gbase[tid * D::kDpt + i] = goff[i] - start[i];
  }
  // payloads out of their striped slots before the reorder overwrites them
  int32_t val[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j)
    val[j] = kFirst ? (int32_t)(wbase + 64 * j + lane - r.seg_begin)
                    : s_vals[wave * 64 * kItems + 64 * j + lane];
  __syncthreads();

  // Reorder the tile in LDS by (digit, input order).
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    if ((wbase + 64 * j + lane) < r.end) {
      const uint32_t dj = (key[j] >> shift) & (D::kR - 1);
      const uint32_t wo =
          kPack ? (s_whist[wave][dj >> 1] >> ((dj & 1) << 4)) & 0xffffu
                : s_whist[wave][dj];
      const uint32_t lp = s_start[dj] + wo + rank[j];
      s_keys[lp] = key[j];
      s_vals[lp] = val[j];
    }
  }
  __syncthreads();

  // Write out: consecutive threads take consecutive tile positions, so each
  // digit run goes to consecutive addresses.  Plain stores: the line at each
  // end of a run is shared with a neighbouring tile's run, and the two
  // halves merge in L2 (non-temporal stores took pass 1 from 62 to 187 us).
  // Positions outside the segment cannot occur when the counts are right;
  // the guard keeps a corrupted count from writing outside it.
  const uint32_t tile_n = r.end - r.begin;
  for (uint32_t p = tid; p < tile_n; p += kSortThreads) {
    const uint32_t k = s_keys[p];
    const uint32_t g = s_gbase[(k >> shift) & (D::kR - 1)] + p;
    if (g >= r.seg_begin && g < r.seg_end) {
      keys_out[g] = k;
      vals_out[g] = s_vals[p];
    }
  }
}

template <int kBits, int kItems>
void sort_pass(int pass, const uint32_t *keys_in, const int32_t *vals_in,
               uint32_t *keys_out, int32_t *vals_out, uint32_t n,
               uint32_t tiles, uint32_t *cnt, uint32_t *off, bool counted,
               Ctl *ctl, hipStream_t s, const SortSegs &sg, bool segmented,
               uint32_t num_segs, int key_bits) {
  constexpr int kR = Digits<kBits>::kR;
  const uint32_t shift = (uint32_t)pass * kBits;
  uint32_t *hist = ctl->hist + (size_t)pass * kR;
  // a digit of at most kFewBits significant bits: few values, shared
  // counters
  const bool few = key_bits - (int)shift <= kFewBits;
  if (!counted && few)
    hipLaunchKernelGGL((k_count<kBits, true, kItems>), dim3(tiles),
                       dim3(kSortThreads), 0, s, keys_in, n, shift, cnt, sg);
  else if (!counted)
    hipLaunchKernelGGL((k_count<kBits, false, kItems>), dim3(tiles),
                       dim3(kSortThreads), 0, s, keys_in, n, shift, cnt, sg);
  // the digit totals the scatter scans into bases (null: off is complete)
  const uint32_t *bases_from = nullptr;
  // a small single mesh: the scatter reads the counts themselves
  const int small = !segmented && tiles <= kSmallScanTiles;
  const uint32_t *offs = small ? cnt : off;
  if (segmented) {
    hipLaunchKernelGGL(k_seg_scan<kBits>, dim3(num_segs), dim3(kSortThreads),
                       0, s, cnt, off, tiles, sg, num_segs);
  } else if (!small) {
    bases_from = hist;
    if (tiles <= kScanSmall / 4)
      hipLaunchKernelGGL((k_scan_columns<kBits, 64>), dim3(kR), dim3(64), 0,
                         s, cnt, off, tiles, hist);
    else if (tiles <= kScanSmall)
      hipLaunchKernelGGL((k_scan_columns<kBits, 256>), dim3(kR), dim3(256), 0,
                         s, cnt, off, tiles, hist);
    else
      hipLaunchKernelGGL((k_scan_columns<kBits, kScanThreads>), dim3(kR),
                         dim3(kScanThreads), 0, s, cnt, off, tiles, hist);
  }
  if (pass == 0)
    hipLaunchKernelGGL((k_scatter<kBits, true, kItems>), dim3(tiles),
                       dim3(kSortThreads), 0, s, keys_in,
                       (const int32_t *)nullptr, keys_out, vals_out, n, shift,
                       offs, bases_from, small, sg);
  else
    hipLaunchKernelGGL((k_scatter<kBits, false, kItems>), dim3(tiles),
                       dim3(kSortThreads), 0, s, keys_in, vals_in, keys_out,
                       vals_out, n, shift, offs, bases_from, small, sg);
}

__global__ void k_single_key(const uint32_t *keys, uint32_t *sorted,
                             int32_t *perm) {
  sorted[0] = keys[0];
  perm[0] = 0;
}

void launch_single_key(const uint32_t *keys, uint32_t *sorted, int32_t *perm,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_single_key, dim3(1), dim3(1), 0, s, keys, sorted, perm);
}

void launch_sort_pass(int bits, int pass, const uint32_t *keys_in,
                      const int32_t *vals_in, uint32_t *keys_out,
                      int32_t *vals_out, uint32_t n, uint32_t tiles,
                      uint32_t *cnt, uint32_t *off, bool counted, Ctl *ctl,
                      hipStream_t s, const SortSegs *segs,
                      uint32_t num_segs, int key_bits) {
  if (tiles == 0) return;
  const SortSegs one{};
  const SortSegs &sg = segs ? *segs : one;
  // a batched sort keeps kSortTile tiles; a single mesh's tile follows n
  const bool small = !segs && sort_items(n) == kSmallSortItems;
#define LBVH_SORT_PASS(B, I)                                                 \
  sort_pass<B, I>(pass, keys_in, vals_in, keys_out, vals_out, n, tiles, cnt, \
                  off, counted, ctl, s, sg, segs != nullptr, num_segs,       \
                  key_bits)
  if (bits == kWideBits) {
    if (small) LBVH_SORT_PASS(kWideBits, kSmallSortItems);
    else LBVH_SORT_PASS(kWideBits, kSortItems);
  } else {
    if (small) LBVH_SORT_PASS(kRadixBits, kSmallSortItems);
    else LBVH_SORT_PASS(kRadixBits, kSortItems);
  }
#undef LBVH_SORT_PASS
}

}  // namespace lbvh
