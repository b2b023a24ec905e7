// device_util.h -- device helpers shared by the kernel files.
#pragma once

#include "lbvh_internal.h"

namespace lbvh {

// XCD-aware tile id.  MI355X hands workgroups to its 8 XCDs round-robin by
// id (block b runs on XCD b % 8), and each XCD has its own L2.  This
// bijection on [0, tiles) gives each XCD one contiguous run of tiles, so
// data that consecutive tiles share stays in one L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t tiles) {
  const uint32_t q = tiles / kXcds, r = tiles % kXcds, x = b % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

// Keys [begin, end) of a sort tile of tile_keys keys inside its segment
// [seg_begin, seg_end) (one segment [0, n) when segs.tile_seg is null;
// segments use kSortTile).
struct TileRange {
  uint32_t begin, end, seg_begin, seg_end;
};
__device__ __forceinline__ TileRange sort_tile_range(
    uint32_t tile, uint32_t n, const SortSegs &segs,
    uint32_t tile_keys = kSortTile) {
  if (!segs.tile_seg) {
    const uint32_t b = tile * tile_keys;
    return {b, min(b + tile_keys, n), 0u, n};
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

// Lanes of the wave that are `valid` and hold the calling lane's digit d
// (kBits bits), found with one ballot per bit.  Every lane of the wave must
// call it.  Per bit, each half of the peer mask keeps the lanes whose vote
// equals the lane's own bit: peers & ~(vote ^ s), s = the bit spread over
// the word, one v_bitop3 per half (truth table 0x84 = S1 & ~(S0 ^ S2),
// symmetric in S0 and S2).  The loop has no exit before kBits: a uniform
// early exit (to skip digit bits that are zero in every key) measured
// slower than the ballots it saves, as it keeps the items' matches from
// interleaving.
template <int kBits>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  const uint64_t v0 = __ballot(valid);
  uint32_t lo = (uint32_t)v0, hi = (uint32_t)(v0 >> 32);
#pragma unroll
  for (int b = 0; b < kBits; ++b) {
    const uint32_t s = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
    const uint64_t vote = __ballot(s != 0u);
    lo = __builtin_amdgcn_bitop3_b32((uint32_t)vote, lo, s, 0x84);
    hi = __builtin_amdgcn_bitop3_b32((uint32_t)(vote >> 32), hi, s, 0x84);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Non-temporal 16-byte loads and stores (the `nt` bit): for streams that
// are read or written once, so that they do not displace lines that are
// read again.
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef int nt_i4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_nt(const float4 *p) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_nt(float4 *p, float4 v) {
  __builtin_nontemporal_store(nt_f4{v.x, v.y, v.z, v.w},
                              reinterpret_cast<nt_f4 *>(p));
}
__device__ __forceinline__ void store_nt(int4 *p, int4 v) {
  __builtin_nontemporal_store(nt_i4{v.x, v.y, v.z, v.w},
                              reinterpret_cast<nt_i4 *>(p));
}
typedef unsigned int nt_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 load_nt(const uint4 *p) {
  const nt_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u4 *>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 load_nt(const int4 *p) {
  const nt_i4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_i4 *>(p));
  return make_int4(v.x, v.y, v.z, v.w);
}

// Triangles in chunks of kMapThreads: a chunk's 3 * kMapThreads float4 are
// loaded contiguously (each load instruction covers 4 KB, where a triangle
// per lane would stride 48 bytes between lanes), staged in LDS, and thread t
// then reads triangle t of the chunk (staged_box).  A caller loads the next
// chunk into registers before it uses the current one, so those loads are in
// flight meanwhile.  Block size kMapThreads.
struct TriChunks {
  const float4 *tri;
  size_t q_end;  // 3 x (the first triangle past the range)
  float4 r[3];
  // the chunk of triangles [first, first + kMapThreads)
  __device__ void load(uint32_t first) {
    const size_t q0 = 3 * (size_t)first + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const size_t q = q0 + (size_t)j * kMapThreads;
      r[j] = q < q_end ? load_nt(tri + q) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  __device__ void stage(float4 *s_tri) const {
#pragma unroll
    for (int j = 0; j < 3; ++j) s_tri[threadIdx.x + j * kMapThreads] = r[j];
  }
};

// Box of triangle t of a staged chunk: min then max of the three vertices in
// vertex order (aabb-z-idx.js:186-187).
__device__ __forceinline__ void staged_box(const float4 *s_tri, uint32_t t,
                                           float mn[3], float mx[3]) {
  const float4 a = s_tri[3 * t], b = s_tri[3 * t + 1], d = s_tri[3 * t + 2];
  mn[0] = fminf(a.x, fminf(b.x, d.x));
  mn[1] = fminf(a.y, fminf(b.y, d.y));
  mn[2] = fminf(a.z, fminf(b.z, d.z));
  mx[0] = fmaxf(a.x, fmaxf(b.x, d.x));
  mx[1] = fmaxf(a.y, fmaxf(b.y, d.y));
  mx[2] = fmaxf(a.z, fmaxf(b.z, d.z));
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

// ---- Morton keys (aabb-z-idx.js:192-208) ------------------------------------
//
// Shared by the single-mesh and the batched kernels so that both produce the
// same key bits.  Every fp32 operation is an explicitly rounded intrinsic
// and rounds once, as in WGSL.

__device__ __forceinline__ uint32_t split_3(uint32_t u) {
  uint32_t x = u;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// WGSL u32(f32): truncate toward zero, saturate to [0, 2^32-1]; NaN -> 0.
__device__ __forceinline__ uint32_t f32_to_u32_sat(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

// The key of a box centre: rel = (cen - min) / (max - min), divided
// correctly rounded (ieee) or as d * fl(1 / ext), then u32(rel * 1023.99f)
// per axis, interleaved.
__device__ __forceinline__ uint32_t morton_key(const float ce[3],
                                               const float bmin[3],
                                               const float ext[3],
                                               const float rcp[3], bool ieee) {
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float d = __fsub_rn(ce[k], bmin[k]);
    const float rel = ieee ? __fdiv_rn(d, ext[k]) : __fmul_rn(d, rcp[k]);
    q[k] = f32_to_u32_sat(__fmul_rn(rel, 1023.99f));
  }
  return split_3(q[0]) | (split_3(q[1]) << 1) | (split_3(q[2]) << 2);
}

}  // namespace lbvh
