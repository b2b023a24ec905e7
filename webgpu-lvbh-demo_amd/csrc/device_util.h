// device_util.h -- device helpers shared by the kernel files.
#pragma once

#include "lbvh_internal.h"

namespace lbvh {

// Adds one digit per pass to the block's LDS histograms.  A wave whose active
// lanes all hold the same digit (common for the high digits of coherent
// meshes) adds its lane count once instead of contending on one LDS word.
__device__ __forceinline__ void hist_add(uint32_t (*s_hist)[kRadix],
                                         uint32_t key) {
  const uint64_t active = __ballot(1);
  const int leader = __ffsll((unsigned long long)active) - 1;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int p = 0; p < kSortPasses; ++p) {
    const uint32_t d = (key >> (kRadixBits * p)) & (kRadix - 1);
    const uint32_t d0 = __shfl(d, leader, 64);
    const uint64_t same = __ballot(d == d0);
    if (same == active) {
      if (lane == leader) atomicAdd(&s_hist[p][d0], (uint32_t)__popcll(active));
    } else {
      atomicAdd(&s_hist[p][d], 1u);
    }
  }
}

// XCD-aware tile id.  MI355X hands workgroups to its 8 XCDs round-robin by
// id (block b runs on XCD b % 8), and each XCD has its own L2.  This
// bijection on [0, tiles) gives each XCD one contiguous run of tiles, so
// data that consecutive tiles share stays in one L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t tiles) {
  const uint32_t q = tiles / kXcds, r = tiles % kXcds, x = b % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

// Keys [begin, end) of a sort tile inside its segment [seg_begin, seg_end)
// (one segment [0, n) when segs.tile_seg is null).
struct TileRange {
  uint32_t begin, end, seg_begin, seg_end;
};
__device__ __forceinline__ TileRange sort_tile_range(uint32_t tile, uint32_t n,
                                                     const SortSegs &segs) {
  if (!segs.tile_seg) {
    const uint32_t b = tile * (uint32_t)kSortTile;
    return {b, min(b + (uint32_t)kSortTile, n), 0u, n};
  }
  const uint32_t sg = segs.tile_seg[tile];
  const uint32_t sb = segs.seg_off[sg], se = segs.seg_off[sg + 1];
  const uint32_t b = sb + (tile - segs.seg_tile0[sg]) * (uint32_t)kSortTile;
  return {b, min(b + (uint32_t)kSortTile, se), sb, se};
}

// Adds one digit to a block's LDS histogram; a wave whose active lanes all
// hold the same digit adds its lane count once.
__device__ __forceinline__ void digit_add(uint32_t *s_hist, uint32_t d) {
  const uint64_t active = __ballot(1);
  const int leader = __ffsll((unsigned long long)active) - 1;
  const uint32_t d0 = __shfl(d, leader, 64);
  if (__ballot(d == d0) == active) {
    if ((int)(threadIdx.x & 63) == leader)
      atomicAdd(&s_hist[d0], (uint32_t)__popcll(active));
  } else {
    atomicAdd(&s_hist[d], 1u);
  }
}

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Exclusive prefix sum of one value per thread over the whole block (block
// size a multiple of 64, at most 1024 threads).  All threads must call it.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v,
                                                         uint32_t *s_wsum) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_wsum[wave] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (uint32_t w = 0; w < wave; ++w) pre += s_wsum[w];
  __syncthreads();
  return pre + x - v;
}

}  // namespace lbvh
