// k_bounds_morton.hip -- scene-bounds reduction and leaf AABB / Morton keys.
//
// Reference: scripts/bvh/kernels/aabb-z-idx.js (WGSL compute_aabb_z_idx
// :177-195, morton_code :197-199, split_3 :202-208) and the caller-side
// bounds of scripts/obj/parse-obj.js:10-53.
//
// Both kernels are HBM-bound streams over the 48-byte triangles (three
// 16-byte loads per triangle).  Morton arithmetic uses the explicitly
// rounded intrinsics so that every fp32 operation rounds once, as in WGSL.
#include "device_util.h"

namespace lbvh {

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-wide min/max of 6 values (up to 1,024 threads); result valid in
// thread 0.
__device__ void block_minmax(float v[6]) {
  __shared__ float s[16][6];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v[c] = wave_min(v[c]);
    v[3 + c] = wave_max(v[3 + c]);
  }
  if (lane == 0)
    for (int c = 0; c < 6; ++c) s[wave][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      for (int c = 0; c < 3; ++c) {
        v[c] = fminf(v[c], s[w][c]);
        v[3 + c] = fmaxf(v[3 + c], s[w][3 + c]);
      }
  }
  __syncthreads();
}

// Scene bounds: fp32 min/max over all vertex coordinates.  Each block
// reduces a share of the triangles into partials[2*b, 2*b+1]; k_bounds_final
// (one block, launched next) folds the partials into ctl->bounds.  No atomics
// and no fences: a last-block-finishes counter costs one same-address
// device-scope atomic (with a cache writeback for its release) per block,
// which on MI355X serialises across the XCDs.
// cen (x, y, z arrays of cen_stride(n) floats, or null): also the centre of
// each triangle's box, (max + min) * .5f as aabb-z-idx.js:191 computes it, so
// the Morton pass reads 12 bytes per triangle instead of the 48-byte
// triangle.  A block takes chunks of kMapThreads triangles (TriChunks); a
// chunk's centres go out through LDS, one component per wave, 16 bytes per
// lane.
__global__ __launch_bounds__(kMapThreads) void k_bounds(
    const float4 *__restrict__ tri, uint32_t n, float4 *__restrict__ partials,
    Ctl *ctl, float *__restrict__ cen) {
  static_assert(kMapThreads == 256, "a wave stores one component of a chunk");
  __shared__ float4 s_tri[3 * kMapThreads];
  __shared__ float s_cen[3][kMapThreads];
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  const uint32_t t = threadIdx.x;
  const uint32_t chunks = (n + kMapThreads - 1) / kMapThreads;
  const size_t cs = cen_stride(n);
  TriChunks tc{tri, 3 * (size_t)n, {}};
  if (blockIdx.x < chunks) tc.load(blockIdx.x * kMapThreads);
  for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
    tc.stage(s_tri);
    if (c + gridDim.x < chunks) tc.load((c + gridDim.x) * kMapThreads);
    __syncthreads();
    const uint32_t i = c * kMapThreads + t;
    if (i < n) {
      float mn[3], mx[3];
      staged_box(s_tri, t, mn, mx);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        v[k] = fminf(v[k], mn[k]);
        v[3 + k] = fmaxf(v[3 + k], mx[k]);
        s_cen[k][t] = __fmul_rn(__fadd_rn(mx[k], mn[k]), 0.5f);
      }
    }
    __syncthreads();  // s_tri is refilled for the next chunk; s_cen is full
    if (cen && t < 3 * 64) {
      const uint32_t k = t >> 6, q = (t & 63) * 4;  // component, chunk slot
      const uint32_t i0 = c * kMapThreads + q;
      float *dst = cen + k * cs + i0;
      if (i0 + 3 < n) {
        store_nt(reinterpret_cast<float4 *>(dst),
                 make_float4(s_cen[k][q], s_cen[k][q + 1], s_cen[k][q + 2],
                             s_cen[k][q + 3]));
      } else {
        for (uint32_t e = 0; e < 4 && i0 + e < n; ++e) dst[e] = s_cen[k][q + e];
      }
    }
  }
  block_minmax(v);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = make_float4(v[0], v[1], v[2], 0.0f);
    partials[2 * blockIdx.x + 1] = make_float4(v[3], v[4], v[5], 0.0f);
  }
}

// Folds the bounds partials of `blocks` blocks (min.xyz, max.xyz in w of
// thread 0).  Every thread of the block must call it.
__device__ void fold_partials(const float4 *__restrict__ partials,
                              uint32_t blocks, float w[6]) {
  w[0] = w[1] = w[2] = INFINITY;
  w[3] = w[4] = w[5] = -INFINITY;
  for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
    const float4 lo = partials[2 * b], hi = partials[2 * b + 1];
    w[0] = fminf(w[0], lo.x); w[1] = fminf(w[1], lo.y); w[2] = fminf(w[2], lo.z);
    w[3] = fmaxf(w[3], hi.x); w[4] = fmaxf(w[4], hi.y); w[5] = fmaxf(w[5], hi.z);
  }
  block_minmax(w);
}

// One block of kFoldThreads: two partials per thread at the largest grid,
// so the fold waits on one round of loads.
constexpr int kFoldThreads = 1024;
__global__ __launch_bounds__(kFoldThreads) void k_bounds_final(
    const float4 *__restrict__ partials, uint32_t blocks, Ctl *ctl) {
  float w[6];
  fold_partials(partials, blocks, w);
  if (threadIdx.x == 0) {
    ctl->bounds[0] = w[0]; ctl->bounds[1] = w[1]; ctl->bounds[2] = w[2];
    ctl->bounds[3] = 0.0f;
    ctl->bounds[4] = w[3]; ctl->bounds[5] = w[4]; ctl->bounds[6] = w[5];
    ctl->bounds[7] = 0.0f;
  }
}

// compute_aabb_z_idx (aabb-z-idx.js:177-195) over a grid-stride share:
//   box  = min/max of the three vertices            (:186-187)
//   cen  = (max + min) * .5f                          (:191)
//   rel  = (cen - bmin) / (bmax - bmin)               (:192, per div mode)
//   key  = morton(u32(rel * 1023.99f))                (:194-208)
// Optional outputs: the 32-byte AABB (kBoxes, the per-stage entry point;
// staged in LDS so that each store instruction writes whole lines) and,
// for the fused build, the first sort pass's per-tile digit counts,
// cnt[d * tiles + tile] (saving the sort a read of the keys).  With cnt a
// block covers one sort tile (XCD-aware ids, as the sort's), without it a
// grid-stride share; either way in TriChunks chunks.
template <int kBits, bool kBoxes>
__global__ __launch_bounds__(kMapThreads) void k_aabb_morton(
    const float4 *__restrict__ tri, uint32_t n, MortonParams p,
    int bounds_from_ctl, float4 *__restrict__ aabb,
    uint32_t *__restrict__ keys, Ctl *ctl,
    uint32_t *__restrict__ cnt, uint32_t tile_chunks) {
  constexpr int kR = 1 << kBits;
  static_assert(kSortTile % kMapThreads == 0 &&
                    kSmallSortTile % kMapThreads == 0,
                "tiles are whole chunks");
  __shared__ uint32_t s_hist[kR];
  __shared__ float4 s_tri[3 * kMapThreads];
  __shared__ float4 s_box[kBoxes ? 2 * kMapThreads : 1];
  if (cnt)
    for (int d = threadIdx.x; d < kR; d += kMapThreads) s_hist[d] = 0u;
  float bmin[3], ext[3], rcp[3];
  if (bounds_from_ctl) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bmin[c] = ctl->bounds[c];
      ext[c] = __fsub_rn(ctl->bounds[4 + c], bmin[c]);
      rcp[c] = __fdiv_rn(1.0f, ext[c]);
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bmin[c] = p.bmin[c];
      ext[c] = p.ext[c];
      rcp[c] = p.rcp[c];
    }
  }
  const bool ieee = (p.div_mode == 1);
  const uint32_t t = threadIdx.x;
  const uint32_t chunks = (n + kMapThreads - 1) / kMapThreads;
  uint32_t c0 = blockIdx.x, c1 = chunks, cstep = gridDim.x, tile = 0;
  if (cnt) {  // one sort tile of tile_chunks chunks per block
    tile = xcd_tile(blockIdx.x, gridDim.x);
    c0 = tile * tile_chunks;
    c1 = min(c0 + tile_chunks, chunks);
    cstep = 1;
  }
  TriChunks tc{tri, 3 * (size_t)n, {}};
  if (c0 < c1) tc.load(c0 * kMapThreads);
  for (uint32_t c = c0; c < c1; c += cstep) {
    tc.stage(s_tri);
    if (c + cstep < c1) tc.load((c + cstep) * kMapThreads);
    __syncthreads();
    const uint32_t i = c * kMapThreads + t;
    if (i < n) {
      float mn[3], mx[3];
      staged_box(s_tri, t, mn, mx);
      float ce[3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        ce[k] = __fmul_rn(__fadd_rn(mx[k], mn[k]), 0.5f);
      const uint32_t key = morton_key(ce, bmin, ext, rcp, ieee);
      __builtin_nontemporal_store(key, keys + i);
      if (kBoxes) {
        s_box[2 * t] = make_float4(mn[0], mn[1], mn[2], 0.0f);
        s_box[2 * t + 1] = make_float4(mx[0], mx[1], mx[2], 0.0f);
      }
      if (cnt) digit_add(s_hist, key & (kR - 1));
    }
    __syncthreads();  // s_tri is refilled for the next chunk; s_box is full
    if (kBoxes) {
      // the chunk's boxes, 16 contiguous bytes per lane (read back before
      // the next chunk's boxes are staged, past the next barrier)
      const uint32_t nb = 2 * (min(c * kMapThreads + kMapThreads, n) -
                               c * kMapThreads);
      float4 *o = aabb + 2 * (size_t)c * kMapThreads;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t e = j * kMapThreads + t;
        if (e < nb) store_nt(o + e, s_box[e]);
      }
    }
  }
  if (cnt) {
    __syncthreads();  // also when the block had no chunk
    for (int d = threadIdx.x; d < kR; d += kMapThreads)
      cnt[(size_t)d * gridDim.x + tile] = s_hist[d];
  }
}

// compute_aabb_z_idx from the box centres k_bounds wrote (aabb-z-idx.js
// :192-195 onward; the centre of :191 is already rounded the same way), plus
// the first sort pass's per-tile digit counts.  One block per sort tile of
// kTileKeys keys, as k_aabb_morton with cnt.  Bounds from ctl, or (parts != 0) folded by every
// block from the bounds partials as k_bounds_final folds them, which saves
// small builds that launch.
template <int kBits, int kTileKeys>
__global__ __launch_bounds__(kMapThreads) void k_morton_cen(
    const float *__restrict__ cen, uint32_t n, int div_mode,
    uint32_t *__restrict__ keys, const Ctl *ctl, uint32_t *__restrict__ cnt,
    const float4 *__restrict__ partials, uint32_t parts) {
  constexpr int kR = 1 << kBits;
  __shared__ uint32_t s_hist[kR];
  __shared__ float s_b[8];
  for (int d = threadIdx.x; d < kR; d += kMapThreads) s_hist[d] = 0u;
  if (parts) {
    float w[6];
    fold_partials(partials, parts, w);
    if (threadIdx.x == 0) {
      s_b[0] = w[0]; s_b[1] = w[1]; s_b[2] = w[2];
      s_b[4] = w[3]; s_b[5] = w[4]; s_b[6] = w[5];
    }
  } else if (threadIdx.x < 8) {
    s_b[threadIdx.x] = ctl->bounds[threadIdx.x];
  }
  __syncthreads();
  float bmin[3], ext[3], rcp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    bmin[c] = s_b[c];
    ext[c] = __fsub_rn(s_b[4 + c], bmin[c]);
    rcp[c] = __fdiv_rn(1.0f, ext[c]);
  }
  const bool ieee = (div_mode == 1);
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t first = tile * (uint32_t)kTileKeys;
  const uint32_t end = min(first + (uint32_t)kTileKeys, n);
  const size_t cs = cen_stride(n);
  const auto key_of = [&](float x, float y, float z) {
    const float ce[3] = {x, y, z};
    return morton_key(ce, bmin, ext, rcp, ieee);
  };
  // All of the thread's centres are loaded before any is used: one
  // iteration's loads per wave in flight leave the kernel latency-bound far
  // below HBM bandwidth.  A whole tile goes 16 bytes per lane: thread t takes
  // triangles first + 1024 g + 4 t .. + 3 of each group g (counting is
  // indifferent to which thread holds which key).
  constexpr int kPer = kTileKeys / kMapThreads;
  static_assert(kPer % 4 == 0, "whole float4 groups per thread");
  if (end - first == (uint32_t)kTileKeys) {
    float4 c[3][kPer / 4];
#pragma unroll
    for (int g = 0; g < kPer / 4; ++g) {
      const size_t i = first + (size_t)g * 4 * kMapThreads + 4 * threadIdx.x;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        c[k][g] = load_nt(reinterpret_cast<const float4 *>(cen + k * cs + i));
    }
#pragma unroll
    for (int g = 0; g < kPer / 4; ++g) {
      const size_t i = first + (size_t)g * 4 * kMapThreads + 4 * threadIdx.x;
      uint4 kq;
      kq.x = key_of(c[0][g].x, c[1][g].x, c[2][g].x);
      kq.y = key_of(c[0][g].y, c[1][g].y, c[2][g].y);
      kq.z = key_of(c[0][g].z, c[1][g].z, c[2][g].z);
      kq.w = key_of(c[0][g].w, c[1][g].w, c[2][g].w);
      // plain stores: the sort's first scatter reads the keys next
      *reinterpret_cast<uint4 *>(keys + i) = kq;
      digit_add(s_hist, kq.x & (kR - 1));
      digit_add(s_hist, kq.y & (kR - 1));
      digit_add(s_hist, kq.z & (kR - 1));
      digit_add(s_hist, kq.w & (kR - 1));
    }
  } else {
    float c[3][kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const uint32_t i = first + threadIdx.x + j * kMapThreads;
#pragma unroll
      for (int k = 0; k < 3; ++k) c[k][j] = i < end ? cen[k * cs + i] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const uint32_t i = first + threadIdx.x + j * kMapThreads;
      if (i >= end) continue;
      const uint32_t key = key_of(c[0][j], c[1][j], c[2][j]);
      __builtin_nontemporal_store(key, keys + i);
      digit_add(s_hist, key & (kR - 1));
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kR; d += kMapThreads)
    cnt[(size_t)d * gridDim.x + tile] = s_hist[d];
}

__global__ void k_copy_bounds(const Ctl *ctl, float *out) {
  const int t = threadIdx.x;
  if (t < 3) out[t] = ctl->bounds[t];
  else if (t < 6) out[t] = ctl->bounds[t + 1];
}

uint32_t launch_bounds(const float4 *tri, uint32_t n, float4 *partials,
                       Ctl *ctl, hipStream_t s, float *cen,
                       uint32_t fold_max) {
  uint32_t blocks = (n + kMapThreads - 1) / kMapThreads;
  if (blocks > (uint32_t)kBoundsBlocks) blocks = kBoundsBlocks;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(kMapThreads), 0, s, tri, n,
                     partials, ctl, cen);
  if (blocks <= fold_max) return blocks;  // the consumer folds the partials
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kFoldThreads), 0, s,
                     partials, blocks, ctl);
  return 0;
}

void launch_aabb_morton(const float4 *tri, uint32_t n, const MortonParams &p,
                        bool bounds_from_ctl, float4 *aabb, uint32_t *keys,
                        Ctl *ctl, hipStream_t s, uint32_t *cnt, int bits) {
  uint32_t blocks = (n + kMapThreads - 1) / kMapThreads;
  if (blocks > (uint32_t)kMortonBlocksMax) blocks = kMortonBlocksMax;
  if (cnt) blocks = sort_tiles(n);
  if (blocks == 0) blocks = 1;
  // chunks of kMapThreads triangles per sort tile
  const uint32_t tile_chunks =
      (uint32_t)(kSortThreads * sort_items(n)) / kMapThreads;
  const bool wide = cnt && bits == kWideBits;  // digits of the counts
#define LBVH_AABB_MORTON(B, X)                                              \
  hipLaunchKernelGGL((k_aabb_morton<B, X>), dim3(blocks), dim3(kMapThreads), \
                     0, s, tri, n, p, bounds_from_ctl ? 1 : 0, aabb, keys,  \
                     ctl, cnt, tile_chunks)
  if (aabb) {
    if (wide) LBVH_AABB_MORTON(kWideBits, true);
    else LBVH_AABB_MORTON(kRadixBits, true);
  } else {
    if (wide) LBVH_AABB_MORTON(kWideBits, false);
    else LBVH_AABB_MORTON(kRadixBits, false);
  }
#undef LBVH_AABB_MORTON
}

void launch_morton_cen(const float *cen, uint32_t n, int div_mode,
                       uint32_t *keys, const Ctl *ctl, uint32_t *cnt, int bits,
                       hipStream_t s, const float4 *partials, uint32_t parts) {
  const uint32_t tiles = sort_tiles(n);
  if (tiles == 0) return;
  const bool small = sort_items(n) == kSmallSortItems;
#define LBVH_MORTON_CEN(B, T)                                              \
  hipLaunchKernelGGL((k_morton_cen<B, T>), dim3(tiles), dim3(kMapThreads), \
                     0, s, cen, n, div_mode, keys, ctl, cnt, partials, parts)
  if (bits == kWideBits) {
    if (small) LBVH_MORTON_CEN(kWideBits, kSmallSortTile);
    else LBVH_MORTON_CEN(kWideBits, kSortTile);
  } else {
    if (small) LBVH_MORTON_CEN(kRadixBits, kSmallSortTile);
    else LBVH_MORTON_CEN(kRadixBits, kSortTile);
  }
#undef LBVH_MORTON_CEN
}

void launch_copy_bounds(const Ctl *ctl, float *out, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_bounds, dim3(1), dim3(64), 0, s, ctl, out);
}

}  // namespace lbvh
