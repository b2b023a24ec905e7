// lbvh_internal.h -- shared definitions of the HIP kernels and the host API.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbvh {

// ---- sizes and launch shapes --------------------------------------------

// Sort digits: 8 bits (4 passes over the 30-bit Morton keys: bits 0-7,
// 8-15, 16-23, 24-29) or, as a tunable, 10 bits (3 passes: 0-9, 10-19,
// 20-29).  The stand-alone sort takes any u32 key: 4 passes of 8 bits, or of
// 10 bits (the last one 2 bits wide) when that width is fixed.
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;  // 256 digit values per pass
constexpr int kWideBits = 10;
constexpr int kWideRadix = 1 << kWideBits;
constexpr int kKeyBits = 30;      // Morton keys
constexpr int kFullKeyBits = 32;  // keys of the stand-alone sort
constexpr int kHistWords = 4 * kWideRadix;  // >= passes x digits of any sort
__host__ __device__ constexpr int sort_passes(int bits,
                                             int key_bits = kKeyBits) {
  return (key_bits + bits - 1) / bits;
}
__host__ __device__ constexpr int sort_radix(int bits) { return 1 << bits; }
static_assert(sort_passes(kRadixBits) * kRadix <= kHistWords, "hist size");
static_assert(sort_passes(kWideBits) * kWideRadix <= kHistWords, "hist size");
static_assert(sort_passes(kRadixBits, kFullKeyBits) * kRadix <= kHistWords,
              "hist size");
static_assert(sort_passes(kWideBits, kFullKeyBits) * kWideRadix <= kHistWords,
              "hist size");
constexpr int kSortThreads = 256;        // 4 wave64 per tile
constexpr int kSortItems = 16;           // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 keys per tile
// A single-mesh sort of at most kSmallSortMax keys uses tiles of
// kSmallSortTile keys (4 per thread): a small sort is a chain of one-tile
// latencies, and a quarter of the serial work per thread shortens each
// (builds of 2^16 / 2^17 / 2^18 triangles 0.092 / 0.105 / 0.115 -> 0.076 /
// 0.087 / 0.100 ms); above 2^20 keys the four times longer count columns
// cost more than the shorter tiles save.
constexpr int kSmallSortItems = 4;
constexpr int kSmallSortTile = kSortThreads * kSmallSortItems;  // 1024 keys
constexpr uint32_t kSmallSortMax = 1u << 20;

constexpr int kTreeTileDefault = 1024;  // leaves per tile of the fused up pass
constexpr int kMaxTreeLevels = 16;      // agglomeration levels (6 at 2^24, T=1024)
constexpr int kTreeFrontier = 128;      // item slots per agglomeration tile
constexpr int kItemInt4 = 3;            // 48-byte items: range, box, gap deltas

constexpr int kMapThreads = 256;  // triangle-parallel kernels
constexpr int kBoundsBlocks = 2048;  // grid of the bounds reduction
constexpr int kMortonBlocksMax = 2048;  // grid-stride Morton (per-stage entry)

// ---- device control block ------------------------------------------------
//
// One per context, zeroed at allocation.  Nothing in it needs a reset
// between operations (or graph replays):
//   fault       sticky bits (kFault*), never set in a correct run; read and
//               cleared by lbvh_synchronize
//   bounds      scene bounds of the last reduction (k_bounds_final)
//   hist        global digit counts of each pass of the current sort (pass
//               p's at hist + p * radix)
constexpr uint32_t kFaultUpFrontier = 4u;  // per-stage up pass: a block left
                                           // more than kUpFront items, or
                                           // the last level left items
constexpr uint32_t kFaultFrontier = 8u;    // agglomeration frontier overflow
constexpr uint32_t kFaultNotATree = 16u;   // per-stage up pass: an index or
                                           // parent out of range, a path
                                           // longer than any tree's, a third
                                           // child at a node, or a root that
                                           // never merged
struct Ctl {
  uint32_t fault;
  uint32_t pad0[7];
  float bounds[8];                       // min.xyz, 0, max.xyz, 0
  uint32_t hist[kHistWords];             // global digit counts per pass
};


// Morton quantisation parameters (uniform block of aabb-z-idx.js:164-169).
struct MortonParams {
  float bmin[3];
  float ext[3];  // max - min
  float rcp[3];  // fl(1 / ext)
  int div_mode;  // lbvh_div_mode
};

// ---- kernel launchers (each .hip file owns its kernels) --------------------

// k_bounds_morton.hip
// Scene bounds into ctl->bounds; cen (3 x cen_stride(n) floats, or null):
// also each triangle's box centre, for launch_morton_cen.  When the reduction leaves
// at most fold_max partials, the folding launch is skipped and the partial
// count is returned (launch_morton_cen folds them); otherwise 0.
uint32_t launch_bounds(const float4 *tri, uint32_t n, float4 *partials,
                       Ctl *ctl, hipStream_t s, float *cen = nullptr, uint32_t fold_max = 0);
// Morton keys + per-tile digit counts (as launch_aabb_morton with cnt) from
// the centres launch_bounds wrote and the bounds in ctl, or folded from
// `parts` bounds partials when parts != 0.
void launch_morton_cen(const float *cen, uint32_t n, int div_mode,
                       uint32_t *keys, const Ctl *ctl, uint32_t *cnt,
                       int bits, hipStream_t s,
                       const float4 *partials = nullptr, uint32_t parts = 0);
// Bounds partials a Morton block folds itself (beyond this, k_bounds_final).
constexpr uint32_t kFoldPartialsMax = 512;
// Box centres: x, y and z arrays of cen_stride(n) floats each (a multiple of
// 64, so every array and every 256-triangle chunk starts 256-byte aligned
// and is read and written 16 bytes per lane).
__host__ __device__ constexpr uint32_t cen_stride(uint32_t n) {
  return (n + 63u) & ~63u;
}
// cnt (sort_radix(bits) x sort_tiles(n) words, or null): also write the
// first sort pass's per-tile digit counts for digits of `bits` bits.
void launch_aabb_morton(const float4 *tri, uint32_t n, const MortonParams &p,
                        bool bounds_from_ctl, float4 *aabb, uint32_t *keys,
                        Ctl *ctl, hipStream_t s, uint32_t *cnt = nullptr, int bits = kRadixBits);
void launch_copy_bounds(const Ctl *ctl, float *out, hipStream_t s);

// k_sort.hip
// Segments of a batched sort (all null: one segment [0, n)).  Tiles of
// kSortTile keys never straddle segments; tile_seg maps each tile to its
// segment, seg_tile0 gives a segment's first tile, seg_off its first key
// (S+1 entries).
struct SortSegs {
  const uint32_t *tile_seg;
  const uint32_t *seg_tile0;
  const uint32_t *seg_off;
};
// One digit pass of `bits`-bit digits (kRadixBits or kWideBits) over
// `tiles` sort tiles (pass 0 takes the position in the segment as payload
// and ignores vals_in).  cnt and off hold sort_radix(bits) x tiles words
// each; counted: cnt already holds this pass's per-tile digit counts (the
// Morton kernels write pass 0's).  A single mesh takes its digit bases from
// the column sums of cnt (into ctl->hist); each segment from its own tiles'
// counts, in one block per segment.  keys_in and keys_out must differ.
// key_bits: significant bits of the keys (a digit with few of them is
// counted with matched lanes).
void launch_sort_pass(int bits, int pass, const uint32_t *keys_in,
                      const int32_t *vals_in, uint32_t *keys_out,
                      int32_t *vals_out, uint32_t n, uint32_t tiles,
                      uint32_t *cnt, uint32_t *off, bool counted, Ctl *ctl,
                      hipStream_t s, const SortSegs *segs = nullptr,
                      uint32_t num_segs = 0, int key_bits = kKeyBits);
// Keys per thread and tiles of a single-mesh sort of n keys (batched sorts
// always use kSortTile: batch_sort_tiles).
inline int sort_items(uint32_t n) {
  return n <= kSmallSortMax ? kSmallSortItems : kSortItems;
}
inline uint32_t sort_tiles(uint32_t n) {
  const uint32_t t = (uint32_t)(kSortThreads * sort_items(n));
  return (n + t - 1) / t;
}
inline uint32_t batch_sort_tiles(uint32_t n) {
  return (n + kSortTile - 1) / kSortTile;
}
// The sort of a single key: perm[0] = 0 and sorted[0] = keys[0] (sorted may
// be keys).  A one-thread kernel, so that a build stays kernels only.
void launch_single_key(const uint32_t *keys, uint32_t *sorted, int32_t *perm,
                       hipStream_t s);

// k_tree.hip
// The per-stage radix tree (parents array) in three launches over 4,096-gap
// tiles; scratch of radix_tree_scratch_bytes(n) (spine records, per-tile
// summaries), nothing in it needs clearing.
uint64_t radix_tree_scratch_bytes(uint32_t n);
void launch_radix_tree(const uint32_t *keys, uint32_t n, int32_t *parents,
                       void *scratch, hipStream_t s);
// Per-stage up pass: the level-1 kernel, then item levels until one block
// holds the root; scratch of up_pass_scratch_bytes(n) (items, then counts).
uint64_t up_pass_scratch_bytes(uint32_t n);
void launch_up_pass(const int32_t *idx, const float4 *aabb,
                    const int32_t *parents, uint32_t n, float4 *nodes,
                    void *scratch, Ctl *ctl, hipStream_t s);
void launch_rearrange(const float4 *tri, const int32_t *idx, uint32_t n,
                      float4 *out, hipStream_t s);
// Segments of one agglomeration level of a batched build (all null: one
// mesh).  Per tile of the level: its segment; per segment: its first tile at
// this level, its first leaf (seg_off, S+1 entries), and (levels >= 1) its
// first tile and tile count at the level below.
struct TreeSegs {
  const uint32_t *tile_seg;
  const uint32_t *seg_tile0;
  const uint32_t *seg_off;
  const uint32_t *prev_tile0;
  const uint32_t *prev_tiles;
};

// Fused leaf gather + radix tree + node boxes by multi-level agglomeration:
// level 0 over tiles of `tile` (256, 512 or 1024) sorted leaves, then item
// levels until one tile remains.  Tiles are numbered level after level; the
// item slots hold kTreeFrontier items of kItemInt4 int4 per tile and the
// counts one word per tile (tree_tiles_total(n, tile) tiles).
//
// Tiles of each level of an n-leaf build (returns the number of levels).
int tree_levels(uint32_t n, int tile, uint32_t *grid);
uint64_t tree_tiles_total(uint32_t n, int tile);
// One agglomeration level (batched builds pass their segments).  items_in /
// count_in: the level below's slots and counts, indexed by its tile number;
// prev_tiles: its tile count (one mesh).
void launch_tree_level(int level, int tile, uint32_t grid, const float4 *tri,
                       const uint32_t *skeys, const int32_t *perm, uint32_t n,
                       float4 *otri, const int4 *items_in,
                       const uint32_t *count_in, uint32_t prev_tiles,
                       int4 *items_out, uint32_t *count_out, float4 *nodes,
                       Ctl *ctl, const TreeSegs &segs, hipStream_t s);

// k_batch.hip
void launch_batch_prologue(uint32_t *seg_ord, uint32_t num_segs,
                           bool init_bounds, Ctl *ctl, hipStream_t s);
// Per-mesh bounds (into seg_ord) and box centres (3 x total floats).
void launch_batch_bounds(const float4 *tri, const SortSegs &segs,
                         uint32_t tiles, uint32_t *seg_ord, float *cen,
                         uint32_t total, hipStream_t s);
// Per-mesh Morton keys and the first sort pass's per-tile digit counts;
// centres from cen when non-null, else from tri.
void launch_batch_morton(const float4 *tri, const float *cen, uint32_t total,
                         const SortSegs &segs, uint32_t tiles,
                         const float *seg_bounds, const uint32_t *seg_ord,
                         int div_mode, uint32_t *keys, uint32_t *cnt0,
                         int bits, hipStream_t s);

// k_raycast.hip
void launch_raycast(const void *nodes, const float *tri, uint32_t n,
                    const float *rays, uint32_t count, float *hits,
                    hipStream_t s);

}  // namespace lbvh

// lbvh_api.hip: a context's device control block, for the server's
// per-request epilogue (which reads and clears its fault word).
struct lbvh_ctx;
lbvh::Ctl *lbvh_context_ctl(lbvh_ctx *ctx);
