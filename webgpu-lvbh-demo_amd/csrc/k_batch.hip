// k_batch.hip -- per-mesh bounds and Morton keys of a batched build.
//
// A batched build treats S independent meshes, concatenated, in one set of
// launches.  Work is cut into sort tiles of kSortTile triangles that never
// straddle a mesh (SortSegs); each block handles one tile.  The arithmetic is
// the single-mesh path's (aabb-z-idx.js:177-208), per mesh.
#include "device_util.h"

namespace lbvh {

// Order-preserving u32 encoding of fp32, so unsigned atomicMin/atomicMax
// compute float min/max exactly.
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// Resets the per-mesh bounds accumulators (when the GPU computes the
// bounds).
__global__ __launch_bounds__(kMapThreads) void k_batch_prologue(
    uint32_t *seg_ord, uint32_t num_segs, int init_bounds) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (init_bounds)
    for (uint64_t i = first; i < 8ull * num_segs; i += stride)
      seg_ord[i] = ((i & 7) < 3) ? 0xFFFFFFFFu : 0u;  // min slots, max slots
}

// Per-mesh fp32 min/max of the vertices, one sort tile per block, and each
// triangle's box centre (3 x total floats: x, then y, then z) for
// k_batch_morton.
__global__ __launch_bounds__(kSortThreads) void k_batch_bounds(
    const float4 *__restrict__ tri, SortSegs segs, uint32_t *seg_ord,
    float *__restrict__ cen, uint32_t total) {
  static_assert(kSortThreads == kMapThreads, "TriChunks block size");
  __shared__ float s[kSortThreads / 64][6];
  __shared__ float4 s_tri[3 * kMapThreads];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t sg = segs.tile_seg[tile];
  const TileRange r = sort_tile_range(tile, 0u, segs);
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  const uint32_t nch = (r.end - r.begin + kMapThreads - 1) / kMapThreads;
  TriChunks tc{tri, 3 * (size_t)r.end, {}};
  if (nch) tc.load(r.begin);
  for (uint32_t k = 0; k < nch; ++k) {
    tc.stage(s_tri);
    if (k + 1 < nch) tc.load(r.begin + (k + 1) * kMapThreads);
    __syncthreads();
    const uint32_t i = r.begin + k * kMapThreads + threadIdx.x;
    if (i < r.end) {
      float mn[3], mx[3];
      staged_box(s_tri, threadIdx.x, mn, mx);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v[c] = fminf(v[c], mn[c]);
        v[3 + c] = fmaxf(v[3 + c], mx[c]);
        __builtin_nontemporal_store(__fmul_rn(__fadd_rn(mx[c], mn[c]), 0.5f),
                                    cen + (size_t)c * cen_stride(total) + i);
      }
    }
    __syncthreads();  // s_tri is refilled for the next chunk
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c] = fminf(v[c], __shfl_xor(v[c], o, 64));
      v[3 + c] = fmaxf(v[3 + c], __shfl_xor(v[3 + c], o, 64));
    }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int c = 0; c < 6; ++c) s[wave][c] = v[c];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    float m = s[0][c];
    for (int w = 1; w < kSortThreads / 64; ++w)
      m = c < 3 ? fminf(m, s[w][c]) : fmaxf(m, s[w][c]);
    if (c < 3) atomicMin(&seg_ord[8 * sg + c], f2ord(m));
    else atomicMax(&seg_ord[8 * sg + c], f2ord(m));
  }
}

// Per-mesh leaf AABB centre -> Morton key, plus the first sort pass's
// per-tile digit counts (tile-major, cnt0[tile * R + d], as the batched
// sort keeps them).  Bounds: host-provided fp32
// (6 per mesh; centres from the triangles) or the GPU accumulators (centres
// from k_batch_bounds, cen non-null).
template <int kBits>
__global__ __launch_bounds__(kSortThreads) void k_batch_morton(
    const float4 *__restrict__ tri, const float *__restrict__ cen,
    uint32_t total, SortSegs segs, const float *__restrict__ seg_bounds,
    const uint32_t *__restrict__ seg_ord, int div_mode,
    uint32_t *__restrict__ keys, uint32_t *__restrict__ cnt0) {
  constexpr int kR = 1 << kBits;
  __shared__ uint32_t s_hist[kR];
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t sg = segs.tile_seg[tile];
  const TileRange r = sort_tile_range(tile, 0u, segs);
  const uint32_t base = r.begin, end = r.end;
  for (int d = threadIdx.x; d < kR; d += kSortThreads) s_hist[d] = 0u;
  float bmin[3], ext[3], rcp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float lo, hi;
    if (seg_bounds) {
      lo = seg_bounds[6 * sg + c];
      hi = seg_bounds[6 * sg + 3 + c];
    } else {
      lo = ord2f(seg_ord[8 * sg + c]);
      hi = ord2f(seg_ord[8 * sg + 3 + c]);
    }
    bmin[c] = lo;
    ext[c] = __fsub_rn(hi, lo);
    rcp[c] = __fdiv_rn(1.0f, ext[c]);
  }
  __syncthreads();
  const bool ieee = (div_mode == 1);
  // Centres from k_batch_bounds, or from the triangles in TriChunks chunks
  // (cen is the same for every block, so the barriers are uniform).
  __shared__ float4 s_tri[3 * kMapThreads];
  const uint32_t nch = (end - base + kMapThreads - 1) / kMapThreads;
  TriChunks tc{tri, 3 * (size_t)end, {}};
  if (!cen && nch) tc.load(base);
  for (uint32_t ch = 0; ch < nch; ++ch) {
    const uint32_t i = base + ch * kMapThreads + threadIdx.x;
    float ce[3] = {0.f, 0.f, 0.f};
    if (cen) {
      if (i < end) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          ce[c] = __builtin_nontemporal_load(cen + (size_t)c * cen_stride(total) + i);
      }
    } else {
      tc.stage(s_tri);
      if (ch + 1 < nch) tc.load(base + (ch + 1) * kMapThreads);
      __syncthreads();
      if (i < end) {
        float mn[3], mx[3];
        staged_box(s_tri, threadIdx.x, mn, mx);
#pragma unroll
        for (int c = 0; c < 3; ++c)
          ce[c] = __fmul_rn(__fadd_rn(mx[c], mn[c]), 0.5f);
      }
      __syncthreads();  // s_tri is refilled for the next chunk
    }
    if (i >= end) continue;
    const uint32_t key = morton_key(ce, bmin, ext, rcp, ieee);
    __builtin_nontemporal_store(key, keys + i);
    digit_add(s_hist, key & (kR - 1));
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kR; d += kSortThreads)
    cnt0[(size_t)tile * kR + d] = s_hist[d];
}

void launch_batch_prologue(uint32_t *seg_ord, uint32_t num_segs,
                           bool init_bounds, Ctl *ctl, hipStream_t s) {
  const uint64_t words = 8ull * num_segs;
  uint32_t blocks = (uint32_t)((words + kMapThreads - 1) / kMapThreads);
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_batch_prologue, dim3(blocks), dim3(kMapThreads), 0, s,
                     seg_ord, num_segs, init_bounds ? 1 : 0);
}

void launch_batch_bounds(const float4 *tri, const SortSegs &segs,
                         uint32_t tiles, uint32_t *seg_ord, float *cen,
                         uint32_t total, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_bounds, dim3(tiles), dim3(kSortThreads), 0, s,
                     tri, segs, seg_ord, cen, total);
}

void launch_batch_morton(const float4 *tri, const float *cen, uint32_t total,
                         const SortSegs &segs, uint32_t tiles,
                         const float *seg_bounds, const uint32_t *seg_ord,
                         int div_mode, uint32_t *keys, uint32_t *cnt0,
                         int bits, hipStream_t s) {
  if (bits == kWideBits)
    hipLaunchKernelGGL(k_batch_morton<kWideBits>, dim3(tiles),
                       dim3(kSortThreads), 0, s, tri, cen, total, segs,
                       seg_bounds, seg_ord, div_mode, keys, cnt0);
  else
    hipLaunchKernelGGL(k_batch_morton<kRadixBits>, dim3(tiles),
                       dim3(kSortThreads), 0, s, tri, cen, total, segs,
                       seg_bounds, seg_ord, div_mode, keys, cnt0);
}

}  // namespace lbvh
