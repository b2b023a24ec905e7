// k_tree.hip -- radix tree, bottom-up node boxes, triangle rearrange.
//
// Reference: scripts/bvh/kernels/radix-tree.js (WGSL :137-238),
// bvh-up-pass.js (WGSL :200-280) and rearrange.js (WGSL :162-169).
//
// Two code paths produce the same BVH bytes:
//   * per-stage kernels mirroring the reference's execute() shims
//     (k_rt_* -> parents array, k_up_pass over parents, k_rearrange);
//   * the agglomeration used by lbvh_build / lbvh_refit / batched builds:
//     k_leaves gathers the triangles into leaf order, computes leaf boxes
//     and builds the Karras tree bottom-up (Apetrei 2014) without a parents
//     array, merging inside LDS wherever both children of a split lie in
//     the block's tile; each tile passes the subtrees it cannot complete
//     (its frontier) to k_items, whose levels merge the frontiers of 32
//     tiles per block until one block completes the root.
//
// Common definitions.  Sorted position i, key k_i.  delta(i, j) = common
// prefix length of the keys at i and j, with equal keys compared by their
// positions (radix-tree.js:172-185):
//     delta = (k_i != k_j) ? clz(k_i ^ k_j) : 32 + clz(i ^ j),  -1 outside.
// A node covering leaves [l, r] was split by its parent either at r (it is
// the parent's left child, Karras id r) or at l - 1 (right child, id l): the
// parent's split is the neighbouring gap with the larger delta.  The root
// [0, N-1] is node 0.  The split-s node's children are numbered s and s+1
// (leaves when their range is a single position), radix-tree.js:226-235.
#include "device_util.h"

namespace lbvh {

// ---------------------------------------------------------------------------
// Per-stage radix tree (compute_radix_tree_pointers, radix-tree.js:137-157,
// compute_child_index :188-238, dif/del :160-185)
// ---------------------------------------------------------------------------

// The parents array by all nearest smaller values.  The Karras tree is
// the min-Cartesian tree of the gap deltas d(g) = delta(g, g+1): the node
// split at gap g covers the leaves between the nearest gaps with a smaller
// delta on either side, L(g) and R(g) (sentinels -1 and n-1, delta -1; two
// equal deltas never face each other across larger ones, so "smaller" is
// unambiguous).  Its parent splits at whichever of L and R has the larger
// delta: at R, the node is the left child and its Karras id is R; at L, the
// right child with id L + 1; the root (both sentinels) is 0.  Its children
// are numbered g and g + 1 (leaves when single), so gap g writes exactly the
// parents slots g or n + g, and g + 1 or n + g + 1 (radix-tree.js:146-156).
//
// Three launches, no waiting between blocks:
//   k_rt_tiles    a block takes 4,096 gaps: deltas from its keys in LDS, L
//                 and R inside 64-gap segments from bit-plane ballots, then
//                 across the tile's segments from per-segment tables of the
//                 last and first gap below each delta value (so a search
//                 costs a few LDS reads); it writes the parents slots of
//                 every gap whose bounds it found.  Gaps with no smaller
//                 delta on one side inside the tile (its left and right
//                 spines, strictly monotone, so at most 64 each) are kept
//                 as records, and the spines give the tile's summary: for
//                 each delta value v, its last and its first gap with a
//                 smaller delta.
//   k_rt_scan     per value v, the summaries of the tiles before (after)
//                 each tile: a prefix max (suffix min) over tiles.
//   k_rt_resolve  each record takes its missing bound from the scans and
//                 writes its slots.
constexpr int kRtThreads = 256;
constexpr int kRtPer = 16;                          // gaps per thread
constexpr int kRtTile = kRtThreads * kRtPer;        // 4096 gaps per tile
constexpr int kRtSegs = kRtTile / 64;               // 64-gap segments
constexpr int kRtRecords = 128;                     // spine gaps per tile
// Deltas are at most 63 (32 + clz of (g xor g + 1) when the keys are equal),
// so six bit planes hold them.  A gap past the end takes 63 too: no lane u
// (a value 0 .. 63) counts it as below u, so it never bounds a real gap, it
// is skipped before it writes anything, and a segment of such gaps has
// minimum 63, below no u.
constexpr int kRtPlanes = 6;                        // delta bits
constexpr int kRtNone = 63;                         // "no gap" delta
constexpr int kRtScanThreads = 1024;  // four tiles per thread at 16M

// Lanes of a segment whose delta is below v, from its bit planes: from the
// high bit down, lanes equal to v so far drop out where their bit differs,
// and those with a 0 where v has a 1 are below it.
__device__ __forceinline__ uint64_t rt_below(const uint64_t *planes, int v) {
  uint64_t lt = 0, eq = ~0ull;
#pragma unroll
  for (int b = kRtPlanes - 1; b >= 0; --b) {
    const uint64_t pb = planes[b];
    const uint64_t vb = 0ull - (uint64_t)((v >> b) & 1);  // v's bit, spread
    lt |= eq & vb & ~pb;
    eq &= ~(pb ^ vb);
  }
  return lt;
}

// Writes gap g's two child slots: its Karras id from its bounds.
__device__ __forceinline__ void rt_write(int32_t *__restrict__ par, uint32_t n,
                                         int64_t g, int64_t L, int dL,
                                         int64_t R, int dR) {
  const int32_t id =
      (dL < 0 && dR < 0) ? 0 : (dL > dR ? (int32_t)(L + 1) : (int32_t)R);
  par[(L + 1 == g) ? (int64_t)n + g : g] = id;
  par[(g + 1 == R) ? (int64_t)n + g + 1 : g + 1] = id;
}

// A deferred gap: {g, v | flags << 8 | (dL + 1) << 16 | (dR + 1) << 24, L, R}
// with flag 1: L missing, 2: R missing.
__global__ __launch_bounds__(kRtThreads) void k_rt_tiles(
    const uint32_t *__restrict__ keys, uint32_t n, int32_t *__restrict__ par,
    int4 *__restrict__ recs, uint32_t *__restrict__ rec_cnt,
    unsigned long long *__restrict__ last,
    unsigned long long *__restrict__ first, uint32_t tiles) {
  __shared__ uint32_t s_key[kRtTile + 1];
  __shared__ uint8_t s_d[kRtTile];
  // [segment][value u]: the segment's last / first gap with a delta below u
  // (any byte where the segment has none)
  __shared__ uint8_t s_last[kRtSegs][64], s_first[kRtSegs][64];
  __shared__ int s_min[kRtSegs];
  __shared__ uint64_t s_segbelow[64];  // [u]: segments whose minimum is below u
  __shared__ int s_lpos[64], s_rpos[64];  // spine gap of each delta value
  __shared__ uint32_t s_nrec;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t base = (int64_t)t * kRtTile;
  const int64_t m = (int64_t)n - 1;  // gaps
  if (t == 0 && tid == 0) {
    par[0] = 0;       // the root's slot: no gap has the root as a child
    par[n - 1] = 0;   // no internal node n-1
    if (n == 1) par[1] = 0;
  }
  for (int i = tid; i <= kRtTile; i += kRtThreads) {
    const int64_t k = base + i;
    s_key[i] = k < (int64_t)n ? keys[k] : 0u;
  }
  if (tid < 64) {
    s_lpos[tid] = -1;
    s_rpos[tid] = -1;
  }
  if (tid == 0) s_nrec = 0;
  __syncthreads();
  // Deltas, and per segment (wave w owns segments 16w .. 16w + 15, lane l
  // gap l of each) its bit planes and its minimum.  With lane u standing for
  // the value u, the planes give the mask of the segment's gaps below u:
  // the tables of its last and first such gap come from it, and it stays in
  // registers for the segment's own gaps (a gap of delta v reads lane v's).
  int dv[kRtPer];
  uint64_t below_u[kRtPer];
#pragma unroll
  for (int j = 0; j < kRtPer; ++j) {
    const int sg = wave * kRtPer + j;
    const int i = sg * 64 + lane;
    const int64_t g = base + i;
    int v = kRtNone;
    if (g < m) {
      const uint32_t a = s_key[i], b = s_key[i + 1];
      v = (a != b) ? (int)__clz(a ^ b)
                   : 32 + (int)__clz((uint32_t)g ^ (uint32_t)(g + 1));
    }
    dv[j] = v;
    s_d[i] = (uint8_t)v;
    uint64_t pl[kRtPlanes];
#pragma unroll
    for (int b = 0; b < kRtPlanes; ++b) pl[b] = __ballot((v >> b) & 1);
    // the segment's minimum from its planes, high bit first (wave-uniform)
    uint64_t cand = ~0ull;
    int mn = 0;
#pragma unroll
    for (int b = kRtPlanes - 1; b >= 0; --b) {
      const uint64_t zero = cand & ~pl[b];
      if (zero) cand = zero;
      else mn |= 1 << b;
    }
    const uint64_t tu = rt_below(pl, lane);
    below_u[j] = tu;
    s_last[sg][lane] = (uint8_t)(tu ? 63 - (int)__clzll((long long)tu) : 0);
    s_first[sg][lane] = (uint8_t)(tu ? (int)__ffsll((long long)tu) - 1 : 0);
    if (lane == 0) s_min[sg] = mn;
  }
  __syncthreads();
  // The segment minima are 64 values, one per lane of a wave: their bit
  // planes give, for each value u, the segments holding a delta below u.
  static_assert(kRtSegs == 64, "one lane per segment");
  if (wave == 0) {
    const int smn = s_min[lane];
    uint64_t sp[kRtPlanes];
#pragma unroll
    for (int b = 0; b < kRtPlanes; ++b) sp[b] = __ballot((smn >> b) & 1);
    s_segbelow[lane] = rt_below(sp, lane);
  }
  __syncthreads();
  const uint64_t below_me = (1ull << lane) - 1;
  const uint64_t above_me = lane == 63 ? 0ull : ~((2ull << lane) - 1);
#pragma unroll 1
  for (int j = 0; j < kRtPer; ++j) {
    const int sg = wave * kRtPer + j;
    const int i = sg * 64 + lane;
    const int64_t g = base + i;
    const int v = dv[j];
    // every lane takes part in the exchange (a gap past the end asks lane 63)
    const uint64_t tu = below_u[j];
    const uint64_t lt =
        (uint64_t)(uint32_t)__shfl((int)(uint32_t)tu, v, 64) |
        ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(tu >> 32), v, 64) << 32);
    if (g >= m) continue;
    const uint64_t ml = lt & below_me, mr = lt & above_me;
    // tile-local bounds, -1: not in the tile
    int li = ml ? sg * 64 + 63 - (int)__clzll((long long)ml) : -1;
    int ri = mr ? sg * 64 + (int)__ffsll((long long)mr) - 1 : -1;
    if (li < 0 || ri < 0) {
      // the nearest other segment whose minimum is below v, then its last
      // (first) gap below v
      const uint64_t segs = s_segbelow[v];
      if (li < 0) {
        const uint64_t sl = segs & ((1ull << sg) - 1);
        if (sl) {
          const int s2 = 63 - (int)__clzll((long long)sl);
          li = s2 * 64 + (int)s_last[s2][v];
        }
      }
      if (ri < 0) {
        const uint64_t sr = sg == 63 ? 0ull : segs & ~((2ull << sg) - 1);
        if (sr) {
          const int s2 = (int)__ffsll((long long)sr) - 1;
          ri = s2 * 64 + (int)s_first[s2][v];
        }
      }
    }
    const int dL = li >= 0 ? (int)s_d[li] : -1;
    const int dR = ri >= 0 ? (int)s_d[ri] : -1;
    if (li >= 0 && ri >= 0) {
      rt_write(par, n, g, base + li, dL, base + ri, dR);
      continue;
    }
    // a spine gap: its value is unique on each spine it is on
    if (li < 0) s_lpos[v] = i;
    if (ri < 0) s_rpos[v] = i;
    const uint32_t r = atomicAdd(&s_nrec, 1u);
    if (r < (uint32_t)kRtRecords) {
      const int flags = (li < 0 ? 1 : 0) | (ri < 0 ? 2 : 0);
      recs[(size_t)t * kRtRecords + r] =
          make_int4((int)g, v | (flags << 8) | ((dL + 1) << 16) | ((dR + 1) << 24),
                    li >= 0 ? (int)(base + li) : 0, ri >= 0 ? (int)(base + ri) : 0);
    }
  }
  __syncthreads();
  // The tile's summary per delta value v (lane v): its last gap with a
  // smaller delta is the right-spine gap with the largest value below v,
  // its first such gap the left-spine gap with the largest value below v.
  if (wave == 0) {
    const uint64_t rv = __ballot(s_rpos[lane] >= 0);
    const uint64_t lv = __ballot(s_lpos[lane] >= 0);
    const uint64_t mr = rv & below_me, ml = lv & below_me;
    unsigned long long e_last = 0ull, e_first = ~0ull;
    if (mr) {
      const int u = 63 - (int)__clzll((long long)mr);
      e_last = ((unsigned long long)(base + s_rpos[u] + 1) << 8) | (unsigned)u;
    }
    if (ml) {
      const int u = 63 - (int)__clzll((long long)ml);
      e_first = ((unsigned long long)(base + s_lpos[u]) << 8) | (unsigned)u;
    }
    last[(size_t)lane * tiles + t] = e_last;
    first[(size_t)lane * tiles + t] = e_first;
    if (lane == 0) rec_cnt[t] = min(s_nrec, (uint32_t)kRtRecords);
  }
}

// Exclusive scans of the tile summaries, in place: block c < 64, prefix max
// of last[c]; block 64 + c, suffix min of first[c] (tiles in reverse).
__global__ __launch_bounds__(kRtScanThreads) void k_rt_scan(
    unsigned long long *__restrict__ last,
    unsigned long long *__restrict__ first, uint32_t tiles) {
  __shared__ unsigned long long s_w[kRtScanThreads / 64];
  const int c = blockIdx.x & 63;
  const bool suffix = blockIdx.x >= 64;
  unsigned long long *col = (suffix ? first : last) + (size_t)c * tiles;
  const unsigned long long id = suffix ? ~0ull : 0ull;
  const auto op = [suffix](unsigned long long a, unsigned long long b) {
    return suffix ? (a < b ? a : b) : (a > b ? a : b);
  };
  const uint32_t per = (tiles + kRtScanThreads - 1) / kRtScanThreads;
  const uint32_t k0 = threadIdx.x * per, k1 = min(k0 + per, tiles);
  // position q of the scan order -> tile (reversed for the suffix)
  const auto at = [&](uint32_t q) -> unsigned long long & {
    return col[suffix ? tiles - 1 - q : q];
  };
  unsigned long long agg = id;
  for (uint32_t q = k0; q < k1; ++q) agg = op(agg, at(q));
  // exclusive block scan of the per-thread aggregates
  unsigned long long x = agg;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long y = __shfl_up(x, o, 64);
    if (lane >= o) x = op(x, y);
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  unsigned long long pre = id;
  for (int k = 0; k < w; ++k) pre = op(pre, s_w[k]);
  unsigned long long ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = id;
  unsigned long long run = op(pre, ex);
  for (uint32_t q = k0; q < k1; ++q) {
    const unsigned long long v = at(q);
    at(q) = run;
    run = op(run, v);
  }
}

// The spine gaps: missing bounds from the scans (none: the sentinels).
__global__ __launch_bounds__(kRtRecords) void k_rt_resolve(
    int32_t *__restrict__ par, uint32_t n, const int4 *__restrict__ recs,
    const uint32_t *__restrict__ rec_cnt,
    const unsigned long long *__restrict__ last,
    const unsigned long long *__restrict__ first, uint32_t tiles) {
  const uint32_t t = blockIdx.x;
  if (threadIdx.x >= rec_cnt[t]) return;
  const int4 r = recs[(size_t)t * kRtRecords + threadIdx.x];
  const int v = r.y & 0xFF, flags = (r.y >> 8) & 3;
  int dL = ((r.y >> 16) & 0xFF) - 1, dR = ((r.y >> 24) & 0xFF) - 1;
  int64_t L = r.z, R = r.w;
  if (flags & 1) {
    const unsigned long long e = last[(size_t)v * tiles + t];
    if (e == 0ull) {
      L = -1;
      dL = -1;
    } else {
      L = (int64_t)(e >> 8) - 1;
      dL = (int)(e & 0xFF);
    }
  }
  if (flags & 2) {
    const unsigned long long e = first[(size_t)v * tiles + t];
    if (e == ~0ull) {
      R = (int64_t)n - 1;
      dR = -1;
    } else {
      R = (int64_t)(e >> 8);
      dR = (int)(e & 0xFF);
    }
  }
  rt_write(par, n, r.x, L, dL, R, dR);
}

// ---------------------------------------------------------------------------
// Shared leaf / node helpers
// ---------------------------------------------------------------------------

// bvh-up-pass.js:208-210: zero components move off the "unwritten" sentinel.
__device__ __forceinline__ void perturb(float mn[3], float mx[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (mn[c] == 0.0f) mn[c] = -1e-8f;
    if (mx[c] == 0.0f) mx[c] = 1e-8f;
  }
}

// Box merges of quiet values.  For fminf / fmaxf the compiler canonicalises
// every operand it cannot prove quiet (a v_max_f32 x, x, x each: six per
// merge of two boxes, on every step of a climb).  The boxes a climb merges
// are always quiet -- results of fminf / fmaxf, or canonicalised where a
// caller's array is loaded -- and on quiet operands the bare instructions
// give exactly fminf's and fmaxf's results (the other operand for a quiet
// NaN; -0 below +0, and perturb keeps exact zeros out of the boxes anyway).
__device__ __forceinline__ float qmin(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float qmax(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

__device__ __forceinline__ void tri_box(const int4 a, const int4 b,
                                        const int4 c, float mn[3],
                                        float mx[3]) {
  const float ax = __int_as_float(a.x), ay = __int_as_float(a.y),
              az = __int_as_float(a.z);
  const float bx = __int_as_float(b.x), by = __int_as_float(b.y),
              bz = __int_as_float(b.z);
  const float cx = __int_as_float(c.x), cy = __int_as_float(c.y),
              cz = __int_as_float(c.z);
  mn[0] = fminf(ax, fminf(bx, cx)); mx[0] = fmaxf(ax, fmaxf(bx, cx));
  mn[1] = fminf(ay, fminf(by, cy)); mx[1] = fmaxf(ay, fmaxf(by, cy));
  mn[2] = fminf(az, fminf(bz, cz)); mx[2] = fmaxf(az, fmaxf(bz, cz));
}

__device__ __forceinline__ int4 pack3(const float v[3], int w) {
  return make_int4(__float_as_int(v[0]), __float_as_int(v[1]),
                   __float_as_int(v[2]), w);
}

// Node of a single leaf (N = 1): the reference's lone thread finds parent
// slot 0, arrives first and writes only the l_* half (f_1 = 1).
__device__ void write_single_leaf_node(int4 *nodes, const float mn[3],
                                       const float mx[3]) {
  nodes[0] = pack3(mn, -1);
  nodes[1] = pack3(mx, 1);
  nodes[2] = make_int4(0, 0, 0, 0);
  nodes[3] = make_int4(0, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Per-stage up pass (bvh_upward_pass, bvh-up-pass.js:200-280)
// ---------------------------------------------------------------------------

// One thread per leaf climbs through the parents array.  Each child writes
// its half of the parent node (Karras order: a child is the left one when
// the child numbered one above it has the same parent); the first child to
// arrive stops, the second merges the boxes and climbs on.  Unlike the
// reference there is no zero-sentinel spin (bvh-up-pass.js:265-268).
//
// A block takes kUpTile consecutive leaves.  Parents whose id lies in the
// block's range meet in LDS (per-side box slots and an arrival count); a
// subtree that lies inside the range has all its node ids there, so it
// completes on chip.  The subtrees that cannot complete there -- the
// block's maximal complete subtrees, at most two per level of the tree, so
// at most kUpFront -- leave as items, in leaf order, for k_up_items, which
// merges the items of kUpFanIn neighbouring blocks the same way, level
// after level, until one block holds the root.  No block ever waits for
// another: kernel boundaries order the levels.
constexpr int kUpTile = 1024;
constexpr int kUpFront = 128;  // items a block leaves at most (2 per level)
constexpr int kUpFanIn = 32;   // blocks of the level below per item block
static_assert(kUpFanIn < kUpTile && kUpFront < kUpTile,
              "one thread per source block; a round takes new items");

struct UpLds {
  float box[2][6][kUpTile];  // [side][min.xyz max.xyz][parent id - block start]
  int32_t cid[2][kUpTile];   // [side][parent id - block start]: child id
  uint32_t cnt[kUpTile];     // arrivals per parent
  // [side][parent id - block start]: the child's far leaf (its first leaf
  // when it is the left child, its last when the right one), which the
  // second arriver takes as the merged range's other end
  int32_t far[2][kUpTile];
  // the block's parents entries, ids ts .. te (one past the block, for the
  // sibling test at its edge): a climb reads these at every step, and from
  // global memory they made each step a chain of dependent loads
  int32_t par_leaf[kUpTile + 1];  // parents[n + id]
  int32_t par_node[kUpTile + 1];  // parents[id]
};

// parents[n + id] / parents[id] for id in the block's staged range
// [ts, ts + kUpTile].  A climb only asks for ids there: it climbs through
// leaves and nodes of the block's range [ts, te) and reads their successor
// (<= te) and, for a node inside the range, its parent entry.  The index is
// clamped all the same, and the arrays are read from LDS only: a choice
// between LDS and global memory per lane compiles to a generic (flat) load,
// which waits on the vector memory path at every step of the climb.
__device__ __forceinline__ int32_t up_par_leaf(const UpLds &L, int32_t ts,
                                               int32_t id) {
  const int32_t k = id - ts;
  return L.par_leaf[k < 0 ? 0 : (k > kUpTile ? kUpTile : k)];
}
__device__ __forceinline__ int32_t up_par_node(const UpLds &L, int32_t ts,
                                               int32_t id) {
  const int32_t k = id - ts;
  return L.par_node[k < 0 ? 0 : (k > kUpTile ? kUpTile : k)];
}

// Node num (a leaf when leaf) is its parent p's left child when the child
// numbered one above it has the same parent.
__device__ __forceinline__ bool up_is_left(const UpLds &L, uint32_t n,
                                           int32_t ts, int32_t num,
                                           int32_t p) {
  const int32_t sib = num + 1;
  return (sib <= (int32_t)n - 1 && up_par_leaf(L, ts, sib) == p) ||
         (sib <= (int32_t)n - 2 && up_par_node(L, ts, sib) == p);
}

// An item: a complete subtree that has still to arrive at its parent.
//   [0] child id as the parent lists it, first leaf, last leaf, parent id
//   [1] box min.xyz, side at the parent (0: left child, 1: right child)
//   [2] box max.xyz, 0
__device__ __forceinline__ void up_put_item(int4 *slot, int self, int32_t lo,
                                            int32_t hi, int32_t p, int side,
                                            const float mn[3],
                                            const float mx[3]) {
  slot[0] = make_int4(self, lo, hi, p);
  slot[1] = pack3(mn, side);
  slot[2] = pack3(mx, 0);
}

// Writes the block's pending items in leaf order: each pending thread
// passes its subtree's position in the block (0 .. kUpTile-1), a thread
// that is not pending -1.  mark is kUpTile words of free LDS.  count_out
// gets the number written; more than kUpFront, or any at all when `last`
// (the block held the whole tree), sets ctl->fault.  Returns the number of
// pending items to every thread.
__device__ __forceinline__ uint32_t up_emit(uint32_t *mark, uint32_t *s_wsum,
                                        int at, int4 *items,
                                        uint32_t *count_out, Ctl *ctl,
                                        int last, int self, int32_t lo,
                                        int32_t hi, int32_t p, int side,
                                        const float mn[3],
                                        const float mx[3]) {
  const int tid = threadIdx.x;
  mark[tid] = 0u;
  __syncthreads();
  if (at >= 0) mark[at] = 1u;
  __syncthreads();
  const uint32_t v = mark[tid];
  const uint32_t pos = block_exclusive_scan(v, s_wsum);
  mark[tid] = pos;
  // the scan leaves the per-wave sums in s_wsum
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < kUpTile / 64; ++w) total += s_wsum[w];
  if (tid == kUpTile - 1) {
    *count_out = total < (uint32_t)kUpFront ? total : (uint32_t)kUpFront;
    if (total > (uint32_t)kUpFront || (last && total > 0))
      atomicOr(&ctl->fault, kFaultUpFrontier);
  }
  __syncthreads();
  if (at >= 0) {
    const uint32_t q = mark[at];
    if (q < (uint32_t)kUpFront)
      up_put_item(items + 3 * (size_t)q, self, lo, hi, p, side, mn, mx);
  }
  return total;
}

// The number of nodes that got exactly two children in this block, summed
// over the block (kUpTile threads, each passing its own count); the block's
// first thread gets the total.
__device__ __forceinline__ uint32_t up_block_count(uint32_t c,
                                                   uint32_t *s_sum) {
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if (threadIdx.x == 0) *s_sum = 0u;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicAdd(s_sum, c);
  __syncthreads();
  return *s_sum;
}

__global__ __launch_bounds__(kUpTile) void k_up_pass(
    const int32_t *__restrict__ idx, const int4 *__restrict__ aabb,
    const int32_t *__restrict__ par, uint32_t n, int4 *nodes,
    int4 *__restrict__ items, uint32_t *__restrict__ item_cnt,
    uint32_t *__restrict__ done, Ctl *ctl, int last) {
  __shared__ UpLds L;
  __shared__ uint32_t s_wsum[16];
  __shared__ uint32_t s_done;
  const int tid = threadIdx.x;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t ts = (int64_t)tile * kUpTile;
  const int64_t te = min(ts + kUpTile, (int64_t)n);
  const int64_t g = ts + tid;
  // The leaf's box (a gather through the permutation) is loaded while the
  // block stages its parents entries: both round trips before one barrier.
  float mn[3] = {0.f, 0.f, 0.f}, mx[3] = {0.f, 0.f, 0.f};
  // Malformed input is caught without touching the climb: an index out of
  // range here, and a parents array that is not a tree by the count of
  // completed nodes (every internal node gets exactly two children, once,
  // in exactly one block of one level: n - 1 in all, checked by the last
  // level).  A parent out of range or a path past the step cap just ends a
  // climb; every access stays in bounds.
  //
  // The parents words and the permutation entry go out together, clamped
  // into the arrays (n >= 1) so that no branch separates them: the block
  // waits two round trips, those loads and then the gather through the
  // permutation, instead of three.
  const bool in = g < (int64_t)n;
  const int64_t e = ts + kUpTile;  // thread 0 also stages the entry past the tile
  const bool in_e = tid == 0 && e < (int64_t)n;
  const int32_t pl = __builtin_nontemporal_load(par + (in ? (int64_t)n + g : 0));
  const int32_t pn = __builtin_nontemporal_load(par + (in ? g : 0));
  const int32_t pl_e = par[in_e ? (int64_t)n + e : 0], pn_e = par[in_e ? e : 0];
  const uint32_t raw = (uint32_t)__builtin_nontemporal_load(idx + (in ? g : 0));
  if (in) {
    if (raw >= n) atomicOr(&ctl->fault, kFaultNotATree);
    const uint32_t src = raw < n ? raw : 0u;
    const int4 bl = aabb[2 * (size_t)src], bh = aabb[2 * (size_t)src + 1];
    // quiet from here on: the climb merges with qmin / qmax
    mn[0] = __builtin_canonicalizef(__int_as_float(bl.x));
    mn[1] = __builtin_canonicalizef(__int_as_float(bl.y));
    mn[2] = __builtin_canonicalizef(__int_as_float(bl.z));
    mx[0] = __builtin_canonicalizef(__int_as_float(bh.x));
    mx[1] = __builtin_canonicalizef(__int_as_float(bh.y));
    mx[2] = __builtin_canonicalizef(__int_as_float(bh.z));
  }
  L.cnt[tid] = 0u;
  L.par_leaf[tid] = in ? pl : -1;
  L.par_node[tid] = in ? pn : -1;
  if (tid == 0) {
    L.par_leaf[kUpTile] = in_e ? pl_e : -1;
    L.par_node[kUpTile] = in_e ? pn_e : -1;
  }
  __syncthreads();
  // 0: done, 1: first to arrive at a parent in LDS, 2: pending (an item)
  int mode = 0;
  bool leaf = true;
  int64_t num = g;  // leaf position or internal node id
  int32_t p = -1;
  int side = 0;
  int self = 0;  // this climber's child id at p
  int32_t lo = (int32_t)g, hi = (int32_t)g;  // the subtree's leaves
  if (g < (int64_t)n) {
    perturb(mn, mx);
    if (n == 1) {
      write_single_leaf_node(nodes, mn, mx);
    } else {
      if (g == (int64_t)n - 1) {
        const int4 z = make_int4(0, 0, 0, 0);
        int4 *lastn = nodes + 4 * (size_t)(n - 1);
        lastn[0] = z; lastn[1] = z; lastn[2] = z; lastn[3] = z;
      }
      p = L.par_leaf[tid];
      // A valid tree is at most 64 levels deep (each level consumes a bit of
      // the 62-bit position-augmented key); the step cap and the range
      // check keep a malformed parents array from running away.
      for (int step = 0; step < 128; ++step) {
        if (p < 0 || p >= (int32_t)n - 1) break;
        side = up_is_left(L, n, (int32_t)ts, (int32_t)num, p) ? 0 : 1;
        self = leaf ? -(int)(num + 1) : (int)num;
        if (p < ts || p >= te) {  // the parent lies outside the block
          mode = 2;
          break;
        }
        const int slot = (int)(p - ts);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          L.box[side][k][slot] = mn[k];
          L.box[side][3 + k][slot] = mx[k];
        }
        L.cid[side][slot] = self;
        L.far[side][slot] = side == 0 ? lo : hi;
        const uint32_t old = __hip_atomic_fetch_add(
            &L.cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0) {
          mode = 1;
          break;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          mn[k] = qmin(mn[k], L.box[1 - side][k][slot]);
          mx[k] = qmax(mx[k], L.box[1 - side][3 + k][slot]);
        }
        if (side == 0) hi = L.far[1][slot];
        else lo = L.far[0][slot];
        if (p == 0) break;  // root merged
        leaf = false;
        num = p;
        p = up_par_node(L, (int32_t)ts, p);
      }
    }
  }
  __syncthreads();
  // a first arriver whose sibling did not come on chip is pending too (the
  // sibling reaches outside the block); an item writes its half of the
  // parent when it arrives there at a later level
  if (mode == 1 && L.cnt[p - ts] == 1u) mode = 2;
  // The nodes both of whose children arrived on chip, in id order, four
  // lanes per 64-byte node (scattered 16-byte half-node stores from the
  // climbing lanes cost more than the rest of the pass).
  uint32_t completed = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = j * kUpTile + tid;
    const int kk = e >> 2, q = e & 3, sd = q >> 1;
    if (ts + kk < te && L.cnt[kk] == 2u) {
      completed += q == 0 ? 1u : 0u;
      const int c = (q & 1) * 3;  // min or max half of the side's box
      const int w = (q & 1) ? (sd ? 0 : 2) : L.cid[sd][kk];
      store_nt(nodes + 4 * (size_t)(ts + kk) + q,
               make_int4(__float_as_int(L.box[sd][c][kk]),
                         __float_as_int(L.box[sd][c + 1][kk]),
                         __float_as_int(L.box[sd][c + 2][kk]), w));
    }
  }
  const uint32_t total = up_block_count(completed, &s_done);
  if (tid == 0) {
    done[tile] = total;
    if (last && n > 1 && total != n - 1) atomicOr(&ctl->fault, kFaultNotATree);
  }
  // the parents entries are no longer read: their words order the items
  up_emit(reinterpret_cast<uint32_t *>(L.par_node), s_wsum,
          mode == 2 ? (int)(lo - ts) : -1,
          items + 3 * (size_t)tile * kUpFront, item_cnt + tile, ctl, last,
          self, lo, hi, p, side, mn, mx);
}

// One level of items: block b takes the items of blocks [kUpFanIn b,
// kUpFanIn b + kUpFanIn) of the level below, in leaf order.  Those items
// partition a leaf range, and the split of every node still to complete
// inside it is a gap between two adjacent items: arrivals meet in that
// gap's LDS slot (the left child at the gap after its last item, the right
// child at the gap before its first).  Each arrival writes its half of the
// parent node; the second one merges and climbs on.  Subtrees whose sibling
// reaches outside the range are the level's items.
//
// The block sweeps its items in rounds of at most kUpTile, one thread each:
// the items the previous round left pending (read back from the block's own
// output slots; they are the pending subtrees of every leaf so far, at most
// kUpFront), then as many new source items as fit.  A subtree at a round's
// right edge stays pending exactly as at the block's, and meets its sibling
// in a later round.  Typical sources leave a few dozen items each, so a
// level usually takes one round.
struct UpItemLds {
  float box[2][6][kUpTile];   // [side][min.xyz max.xyz][gap]
  int32_t far_item[2][kUpTile];  // [side][gap]: far item index (as far)
  int32_t far_leaf[2][kUpTile];  // [side][gap]: far leaf
  uint32_t cnt[kUpTile];      // arrivals per gap
  uint32_t mark[kUpTile];     // item ordering
  uint32_t base[kUpFanIn + 1];  // first item of each source block
};

__global__ __launch_bounds__(kUpTile) void k_up_items(
    const int4 *__restrict__ items_in, const uint32_t *__restrict__ cnt_in,
    const uint32_t *__restrict__ done_in, uint32_t tiles_in,
    int4 *items_out,  // written, then read back by the next round
    uint32_t *__restrict__ cnt_out, uint32_t *__restrict__ done_out,
    const int32_t *__restrict__ par, uint32_t n, int4 *nodes, Ctl *ctl,
    int last) {
  __shared__ UpItemLds S;
  __shared__ uint32_t s_wsum[16];
  __shared__ uint32_t s_done;
  const int tid = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const uint32_t c0 = tile * kUpFanIn;
  const uint32_t nc = min((uint32_t)kUpFanIn, tiles_in - c0);
  {
    const uint32_t c =
        (uint32_t)tid < nc ? min(cnt_in[c0 + tid], (uint32_t)kUpFront) : 0u;
    const uint32_t ex = block_exclusive_scan(c, s_wsum);
    if (tid <= kUpFanIn) S.base[tid] = ex;  // the total from nc on
  }
  // nodes completed by the source blocks, then by each round here: the last
  // level holds the count of the whole tree
  uint32_t completed = (uint32_t)tid < nc ? done_in[c0 + tid] : 0u;
  int4 *own = items_out + 3 * (size_t)tile * kUpFront;
  __syncthreads();
  const int total = (int)S.base[kUpFanIn];
  int taken = 0;  // source items used by earlier rounds
  int carry = 0;  // items left pending by the previous round
  while (true) {
    const int fresh = min(total - taken, kUpTile - carry);
    const bool fin = taken + fresh == total;
    const int m = carry + fresh;
    S.cnt[tid] = 0u;
    __syncthreads();
    int mode = 0;  // 0: done, 1: first arriver, 2: pending (an item)
    int self = 0, side = 0;
    int32_t p = -1, lo = 0, hi = 0;
    int i0 = tid, i1 = tid;  // the subtree's items
    float mn[3] = {0.f, 0.f, 0.f}, mx[3] = {0.f, 0.f, 0.f};
    if (tid < m) {
      const int4 *it;
      if (tid < carry) {
        it = own + 3 * (size_t)tid;
      } else {
        // source block of item g: the last j < nc with base[j] <= g
        const uint32_t g = (uint32_t)(taken + tid - carry);
        uint32_t j = 0;
#pragma unroll
        for (uint32_t step = kUpFanIn / 2; step > 0; step >>= 1)
          if (j + step < nc && S.base[j + step] <= g) j += step;
        it = items_in + 3 * ((size_t)(c0 + j) * kUpFront + (g - S.base[j]));
      }
      const int4 a = it[0], b = it[1], c = it[2];
      self = a.x; lo = a.y; hi = a.z; p = a.w;
      mn[0] = __int_as_float(b.x); mn[1] = __int_as_float(b.y);
      mn[2] = __int_as_float(b.z); side = b.w;
      mx[0] = __int_as_float(c.x); mx[1] = __int_as_float(c.y);
      mx[2] = __int_as_float(c.z);
      mode = 2;
      for (int step = 0; step < 128; ++step) {
        if (p < 0 || p >= (int32_t)n - 1) {
          mode = 0;
          break;
        }
        const int gap = side == 0 ? i1 : i0 - 1;
        if (gap < 0 || gap >= m - 1) {  // the sibling lies outside the round
          mode = 2;
          break;
        }
        int4 *node = nodes + 4 * (size_t)p;
        node[2 * side] = pack3(mn, self);
        node[2 * side + 1] = pack3(mx, side ? 0 : 2);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          S.box[side][k][gap] = mn[k];
          S.box[side][3 + k][gap] = mx[k];
        }
        S.far_item[side][gap] = side == 0 ? i0 : i1;
        S.far_leaf[side][gap] = side == 0 ? lo : hi;
        const uint32_t old = __hip_atomic_fetch_add(
            &S.cnt[gap], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0) {
          mode = 1;
          break;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          mn[k] = qmin(mn[k], S.box[1 - side][k][gap]);
          mx[k] = qmax(mx[k], S.box[1 - side][3 + k][gap]);
        }
        if (side == 0) {
          i1 = S.far_item[1][gap];
          hi = S.far_leaf[1][gap];
        } else {
          i0 = S.far_item[0][gap];
          lo = S.far_leaf[0][gap];
        }
        if (p == 0) {  // root merged
          mode = 0;
          break;
        }
        // the parent and the two words of the left-child test (up_is_left)
        // depend only on num: three loads in flight together
        const int32_t num = p;
        const int32_t pl = num + 1 <= (int32_t)n - 1 ? par[n + num + 1] : -2;
        const int32_t pn = num + 1 <= (int32_t)n - 2 ? par[num + 1] : -2;
        self = num;
        p = par[num];
        side = (pl == p || pn == p) ? 0 : 1;
      }
    }
    __syncthreads();
    completed += (tid < m - 1 && S.cnt[tid] == 2u) ? 1u : 0u;
    // a first arriver whose sibling never came in this round is pending
    if (mode == 1 && S.cnt[side == 0 ? i1 : i0 - 1] == 1u) mode = 2;
    const uint32_t count =
        up_emit(S.mark, s_wsum, mode == 2 ? i0 : -1, own, cnt_out + tile, ctl,
                last && fin, self, lo, hi, p, side, mn, mx);
    if (fin) break;
    taken += fresh;
    carry = (int)min(count, (uint32_t)kUpFront);
    __syncthreads();  // the emitted items are visible to the whole block
  }
  const uint32_t all = up_block_count(completed, &s_done);
  if (tid == 0) {
    done_out[tile] = all;
    if (last && all != n - 1) atomicOr(&ctl->fault, kFaultNotATree);
  }
}

// rearrange_triangles (rearrange.js:162-169): whole 48-byte records, stored
// through LDS 16 contiguous bytes per lane (a lane's own record would put 48
// bytes between the lanes of each store).
__global__ __launch_bounds__(kMapThreads) void k_rearrange(
    const int4 *__restrict__ tri, const int32_t *__restrict__ idx, uint32_t n,
    int4 *__restrict__ out) {
  __shared__ int4 s_rec[3 * kMapThreads];
  const uint32_t i0 = blockIdx.x * kMapThreads;
  const uint32_t i = i0 + threadIdx.x;
  if (i < n) {
    const uint32_t j = (uint32_t)__builtin_nontemporal_load(idx + i);
    const size_t s = 3 * (size_t)(j < n ? j : 0u);
    s_rec[3 * threadIdx.x] = tri[s];
    s_rec[3 * threadIdx.x + 1] = tri[s + 1];
    s_rec[3 * threadIdx.x + 2] = tri[s + 2];
  }
  __syncthreads();
  const uint32_t nq = 3 * (min(i0 + kMapThreads, n) - i0);
  int4 *o = out + 3 * (size_t)i0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t e = k * kMapThreads + threadIdx.x;
    if (e < nq) store_nt(o + e, s_rec[e]);
  }
}

// ---------------------------------------------------------------------------
// Fused leaf + tree + up pass: multi-level agglomeration
// ---------------------------------------------------------------------------
//
// Level 0 runs over the sorted leaves, one block per tile of T positions:
// each thread gathers its triangle into leaf order (O_TRI), takes the leaf
// box and climbs, merging in LDS whenever both children of a split lie in
// the tile.  A climber stops when its parent split is outside the tile, or
// when it arrives first at a split whose other child never completes in the
// tile.  Those stopped subtrees are the tile's frontier and partition its
// leaves in order.  Each has its parent on one of the two root paths through
// the tile's edge gaps, and a root path of the position-augmented keys has
// at most 64 nodes, so a tile has at most kFrontierMax of them.  They are
// written, in leaf order, as "items" (leaf range, box, the deltas at the
// range's two boundary gaps) into the tile's own kFrontierMax slots, with
// their count.
//
// Level L >= 1 builds the tree over items: one block of kItemTile threads
// takes the frontiers of fan_in(...) consecutive tiles of level L-1, in rounds
// of at most kItemTile items (the previous round's frontier, then the next
// source items), until a single tile holds all of a mesh's remaining items
// and completes its root.  Frontier items are the subtrees hanging off
// root paths, so climbing over them would be a serial chain of merges as
// long as those paths; the item levels compute their nodes in parallel
// instead (k_items).  The level plan depends only on
// the leaf count, so no block waits for another and no counter is shared.
// (A dynamic tile ticket -- one same-address device-scope atomic per tile --
// and a look-back across tiles both serialise on MI355X, where those
// operations cross XCDs: at ~0.1 us per tile they, not the tree work, set
// the stage time.)  Nodes are written exactly once, by the thread that
// merges their children, as soon as their Karras id is known.

constexpr int kFrontierMax = kTreeFrontier;
constexpr int kItemTile = 1024;
// Level L-1 tiles per tile: kFanIn (up to 32 x 128 items, typically a few
// hundred; a round holds at least kItemTile - kFrontierMax new items), or
// all of them, up to kFanInMax, in one last tile.  Depends only on the
// number of level L-1 tiles, so the host's level plan and the kernel agree.
constexpr int kFanIn = 32;
constexpr int kFanInMax = 64;
static_assert(kFanIn <= kFanInMax && kFanInMax < kItemTile &&
                  (kFanInMax & (kFanInMax - 1)) == 0,
              "one thread per source tile; binary search over them");
__host__ __device__ constexpr uint32_t fan_in(uint32_t prev_tiles) {
  return prev_tiles <= (uint32_t)kFanInMax ? (uint32_t)kFanInMax
                                           : (uint32_t)kFanIn;
}

__device__ __forceinline__ int delta_next(uint32_t a, uint32_t b, uint32_t i) {
  return (a != b) ? (int)__clz(a ^ b) : 32 + (int)__clz(i ^ (i + 1u));
}

// Frontier output shared by both level kinds: the stopped subtrees that
// leave the tile, in item order, into the tile's kFrontierMax slots.  An item
// is its leaf range, its box and the deltas at the gaps on either side of
// the range, so the level above needs no sorted keys.  Returns the count
// to every thread.
__device__ __forceinline__ uint32_t emit_frontier(bool out, int l, int r,
                                                  int dl, int dr,
                                                  const float mn[3],
                                                  const float mx[3],
                                                  uint32_t *wsum,
                                                  uint32_t gtile,
                                                  int4 *items_out,
                                                  uint32_t *count_out,
                                                  Ctl *ctl, int tid, int T) {
  const uint32_t rank = block_exclusive_scan(out ? 1u : 0u, wsum);
  uint32_t count = 0;  // the per-wave sums stay in wsum after the scan
  for (int w = 0; w < (T >> 6); ++w) count += wsum[w];
  if (tid == T - 1) {
    count_out[gtile] = count;
    if (count > (uint32_t)kFrontierMax) atomicOr(&ctl->fault, kFaultFrontier);
  }
  if (out && rank < (uint32_t)kFrontierMax) {
    int4 *io = items_out + kItemInt4 * ((size_t)gtile * kFrontierMax + rank);
    io[0] = make_int4(l, r, __float_as_int(mn[0]), __float_as_int(mn[1]));
    io[1] = make_int4(__float_as_int(mn[2]), __float_as_int(mx[0]),
                      __float_as_int(mx[1]), __float_as_int(mx[2]));
    io[2] = make_int4(dl, dr, 0, 0);
  }
  return count;
}

// ---- level 0: leaves ------------------------------------------------------
//
// At level 0 an item is a leaf, so a climber is just its leaf range.  The
// exchange slot of split g (the gap after leaf ts + g) is the node of that
// split in its final layout: each child writes its half (min.xyz | child
// id, max.xyz | far leaf of its range) before the arrival count, and the
// second arrival reads the sibling's half, so the node is complete in LDS
// when the merge happens.  Its Karras id (the merged range's end that faces
// the parent's split) is known one step later and goes to gap_of; after the
// climb the tile writes its nodes in id order, 64 contiguous bytes per four
// lanes (scattered 16-byte stores of whole nodes from the climbing lanes
// cost more than the rest of the level together).
template <int T>
struct LeafLds {
  int dgap[T + 1];    // delta at the gap after leaf ts - 1 + k
  int4 node[T][4];    // split slot -> node being assembled
  uint32_t cnt[T];    // arrivals per split slot
  int gap_of[T];      // node id - ts -> its split slot (-1: not this tile's)
  uint32_t wsum[16];
};

struct Climb {
  int l, r;  // leaf range
  float mn[3], mx[3];
  int open;  // split slot whose node id is not known yet (-1: none)
  int stop;  // 0 root / inactive, 1 pending local arrival, 2 frontier
  int pgap;
};

template <int T>
__device__ __forceinline__ void climb_leaves(Climb &c, LeafLds<T> &L, int n,
                                             int ts, int te) {
  // stop / pgap are set once after the loop: assigned on separate exits, the
  // struct fields were kept in scratch memory
  int stop = 0, pgap = 0;
  while (true) {
    const int dl = L.dgap[c.l - ts];      // gap before the range
    const int dr = L.dgap[c.r - ts + 1];  // gap after the range
    const bool root = (c.l == 0 && c.r == n - 1);
    const bool is_left = dr > dl;
    if (c.open >= 0) {
      const int K = root ? 0 : (is_left ? c.r : c.l);
      L.gap_of[K - ts] = c.open;
      c.open = -1;
    }
    if (root) break;
    const int gap = is_left ? c.r : c.l - 1;  // the parent's split
    if (gap < ts || gap + 2 > te) {           // other child off-tile
      stop = 2;
      break;
    }
    const int slot = gap - ts;
    const int q = is_left ? 0 : 2;
    const int id = (c.l == c.r) ? -(c.l + 1) : (is_left ? c.r : c.l);
    L.node[slot][q] = pack3(c.mn, id);
    L.node[slot][q + 1] = pack3(c.mx, is_left ? c.l : c.r);
    const uint32_t old = __hip_atomic_fetch_add(
        &L.cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == 0) {
      stop = 1;
      pgap = gap;
      break;
    }
    const int4 smn = L.node[slot][2 - q], smx = L.node[slot][3 - q];
    const float sm[3] = {__int_as_float(smn.x), __int_as_float(smn.y),
                         __int_as_float(smn.z)};
    const float sx[3] = {__int_as_float(smx.x), __int_as_float(smx.y),
                         __int_as_float(smx.z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      c.mn[k] = qmin(c.mn[k], sm[k]);
      c.mx[k] = qmax(c.mx[k], sx[k]);
    }
    if (is_left) c.r = smx.w;
    else c.l = smx.w;
    c.open = slot;
  }
  c.stop = stop;
  c.pgap = pgap;
}

// Level 0: tile = T sorted leaf positions.  Gathers O_TRI and the leaf
// boxes, climbs, writes the tile's nodes and its frontier.  With segments
// (batched builds) every segment is an independent mesh: positions, node ids
// and the permutation are relative to its first leaf.
template <int T>
__global__ __launch_bounds__(T) void k_leaves(
    const int4 *__restrict__ tri_all, const uint32_t *__restrict__ skeys_all,
    const int32_t *__restrict__ perm_all, uint32_t n_all,
    int4 *__restrict__ otri_all, int4 *__restrict__ items_out,
    uint32_t *__restrict__ count_out, int4 *nodes_all, Ctl *ctl,
    TreeSegs segs) {
  __shared__ LeafLds<T> L;
  const int tid = threadIdx.x;
  const uint32_t gtile = xcd_tile(blockIdx.x, gridDim.x);
  uint32_t tile0 = 0, leaf_base = 0, N = n_all;
  if (segs.tile_seg) {
    const uint32_t sg = segs.tile_seg[gtile];
    tile0 = segs.seg_tile0[sg];
    leaf_base = segs.seg_off[sg];
    N = segs.seg_off[sg + 1] - leaf_base;
  }
  const uint32_t tile = gtile - tile0;  // tile within the mesh
  const uint32_t *skeys = skeys_all + leaf_base;
  int4 *nodes = nodes_all + 4 * (size_t)leaf_base;
  const int ts = (int)(tile * T);
  const int te = (int)min(tile * T + T, N);
  const int k = ts + tid;
  const bool active = k < te;

  Climb c;
  c.open = -1;
  c.stop = 0;
  c.l = c.r = k;
  L.cnt[tid] = 0u;
  L.gap_of[tid] = -1;
  if (active) {
    // Every load that does not depend on the permutation goes out with it
    // (the keys at the leaf's gaps, clamped to the mesh so no branch
    // separates them): the tile waits two round trips, the permutation and
    // then the gather it names, instead of three.
    const uint32_t pi = (uint32_t)perm_all[leaf_base + k];
    const uint32_t k_here = skeys[k];
    const uint32_t k_next = skeys[(uint32_t)k + 1 < N ? k + 1 : k];
    const uint32_t k_prev = skeys[k > 0 ? k - 1 : k];  // thread 0's gap
    const size_t src = 3 * (size_t)(leaf_base + (pi < N ? pi : 0u));
    // plain loads: neighbouring leaves gather triangles that share lines, and
    // the L2 keeps them (non-temporal loads took the level 0.53 -> 0.65 ms)
    const int4 a = tri_all[src], b = tri_all[src + 1], d = tri_all[src + 2];
    // O_TRI goes out through LDS (the node slots are free until the climb):
    // a lane's own 48-byte row would put 48 bytes between the lanes of each
    // store, where the staged rows are stored 16 contiguous bytes per lane
    int4 *stage = &L.node[0][0];
    stage[3 * tid] = a;
    stage[3 * tid + 1] = b;
    stage[3 * tid + 2] = d;
    tri_box(a, b, d, c.mn, c.mx);
    perturb(c.mn, c.mx);
    // both deltas computed unconditionally and selected, so the compiler
    // does not sink their loads into branches behind the gather
    const int d_next = delta_next(k_here, k_next, (uint32_t)k);
    const int d_prev = delta_next(k_prev, k_here, (uint32_t)(k - 1));
    L.dgap[tid + 1] = ((uint32_t)k + 1 < N) ? d_next : -1;
    if (tid == 0) L.dgap[0] = k > 0 ? d_prev : -1;
  }
  __syncthreads();
  {
    const int4 *stage = &L.node[0][0];
    int4 *o = otri_all + 3 * (size_t)(leaf_base + ts);
    const int nq = 3 * (te - ts);  // the tile's O_TRI int4s
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int e = j * T + tid;
      if (e < nq) store_nt(o + e, stage[e]);
    }
  }
  __syncthreads();  // the node slots are the climb's from here
  if (N == 1) {
    if (k == 0) {
      write_single_leaf_node(nodes, c.mn, c.mx);
      count_out[gtile] = 0u;
    }
    return;
  }
  if ((uint32_t)k == N - 1) {
    const int4 z = make_int4(0, 0, 0, 0);
    int4 *last = nodes + 4 * (size_t)(N - 1);
    last[0] = z; last[1] = z; last[2] = z; last[3] = z;
  }
  if (active) climb_leaves<T>(c, L, (int)N, ts, te);
  __syncthreads();

  // The tile's nodes in id order: int4 e of the tile's node range is
  // quarter e % 4 of node ts + e / 4.  The max halves carried far leaves in
  // .w; the node layout has the constant flags 2 (l_max) and 0 (r_max).
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = j * T + tid;
    const int kk = e >> 2, q = e & 3;
    if (ts + kk < te) {
      const int g = L.gap_of[kk];
      if (g >= 0) {
        int4 v = L.node[g][q];
        if (q == 1) v.w = 2;
        if (q == 3) v.w = 0;
        store_nt(nodes + 4 * (size_t)(ts + kk) + q, v);
      }
    }
  }

  const bool out = active && (c.stop == 2 ||
                              (c.stop == 1 && L.cnt[c.pgap - ts] == 1u));
  emit_frontier(out, c.l, c.r, out ? L.dgap[c.l - ts] : 0,
                out ? L.dgap[c.r - ts + 1] : 0, c.mn, c.mx, L.wsum, gtile,
                items_out, count_out, ctl, tid, T);
}

// ---- levels >= 1: items ---------------------------------------------------

// One item level: tile = the frontiers of fan_in consecutive tiles of the
// level below.  Level L-1's tile g keeps its items in
// items_in[g * kFrontierMax ...] and their number in count_in[g].
//
// The block sweeps them in rounds of m <= kItemTile items: the frontier the
// previous round emitted (read back from the tile's own output slots; it is
// the frontier of every leaf so far, at most kFrontierMax items) followed by
// as many of the next source items as fit.  A round's last gap carries the
// delta to the next round's first leaf, so the nodes it cannot complete
// stay roots, exactly as at the block's own edge.  Typical frontiers are
// small, and a level usually takes a single round.
//
// The tree over the block's m items is the min-Cartesian tree of the deltas
// at the gaps between them: the node split at gap g spans, on each side, up
// to the nearest gap with a smaller delta (equal deltas never face each
// other across a larger-delta stretch, so "smaller" is unambiguous); gap -1
// before the first item and gap m-1 after the last one carry the deltas to
// the leaves outside the block.  A node is complete, and written here, when
// both bounds are found; its parent is the bound with the larger delta.
// So, with one thread per gap and O(log m) LDS steps each instead of a serial
// climb:
//   1. load the items; the gap deltas form a sparse table of range minima
//      (position p = gap + 1) and the item boxes a segment tree;
//   2. each gap finds its nearest smaller delta on either side by binary
//      lifting over the table; a complete node takes its children's boxes
//      from the segment tree and is written at its Karras id;
//   3. the roots of the resulting forest -- complete nodes and items whose
//      parent gap is incomplete or outside the block -- are the frontier,
//      marked at their first item so the emit keeps leaf order.
constexpr int kItemLog = 10;
static_assert(kItemTile == 1 << kItemLog, "sparse table levels");

struct ItemLds {
  signed char dm[kItemLog + 1][kItemTile + 2];  // range minima of gap deltas
  float smn[3][2 * kItemTile], smx[3][2 * kItemTile];  // box segment tree
  int il[kItemTile], ir[kItemTile];  // leaf range of each item
  short qL[kItemTile], qR[kItemTile];  // per gap: nearest smaller positions
  unsigned char complete[kItemTile];  // per gap: node complete in the block
  int root_of[kItemTile];  // per first item: frontier root (gap, T + item)
  uint32_t wsum[16];
  uint32_t pre[kFanInMax + 1];  // first item of each source tile
};

// Box of items [a, b] from the segment tree (leaves at kItemTile + item).
__device__ __forceinline__ void item_range_box(const ItemLds &L, int a, int b,
                                               float mn[3], float mx[3]) {
  mn[0] = mn[1] = mn[2] = INFINITY;
  mx[0] = mx[1] = mx[2] = -INFINITY;
  int lo = a + kItemTile, hi = b + kItemTile + 1;
  while (lo < hi) {
    if (lo & 1) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        mn[c] = qmin(mn[c], L.smn[c][lo]);
        mx[c] = qmax(mx[c], L.smx[c][lo]);
      }
      ++lo;
    }
    if (hi & 1) {
      --hi;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        mn[c] = qmin(mn[c], L.smn[c][hi]);
        mx[c] = qmax(mx[c], L.smx[c][hi]);
      }
    }
    lo >>= 1;
    hi >>= 1;
  }
}

// One round of k_items over m items: items [0, carry) are the tile's own
// output slots (the previous round's frontier), items [carry, m) the source
// items from `taken` on.  Writes the nodes it completes and leaves the
// forest roots in L.root_of (gap, or kItemTile + item) with L.qL / L.qR.
__device__ __forceinline__ void item_round(ItemLds &L, int m, int carry,
                                           int taken, uint32_t src0,
                                           const int4 *__restrict__ items_in,
                                           const int4 *items_out,
                                           uint32_t gtile, uint32_t N,
                                           int4 *nodes) {
  constexpr int T = kItemTile;
  const int tid = threadIdx.x;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    L.smn[c][T + tid] = INFINITY;  // leaves past the items stay neutral
    L.smx[c][T + tid] = -INFINITY;
  }
  L.complete[tid] = 0;
  L.root_of[tid] = -1;
  __syncthreads();

  // 1. items: leaf ranges, boxes (segment-tree leaves), gap deltas
  if (tid < m) {
    const int4 *src;
    if (tid < carry) {
      src = items_out + kItemInt4 * ((size_t)gtile * kFrontierMax + tid);
    } else {
      // the source tile holding item g: the last s with pre[s] <= g
      const int g = taken + tid - carry;
      int s = 0;
#pragma unroll
      for (int step = kFanInMax / 2; step > 0; step >>= 1)
        if ((int)L.pre[s + step] <= g) s += step;
      src = items_in + kItemInt4 * ((size_t)(src0 + s) * kFrontierMax +
                                    (uint32_t)(g - (int)L.pre[s]));
    }
    const int4 a = src[0], b = src[1], g = src[2];
    L.il[tid] = a.x;
    L.ir[tid] = a.y;
    L.smn[0][T + tid] = __int_as_float(a.z);
    L.smn[1][T + tid] = __int_as_float(a.w);
    L.smn[2][T + tid] = __int_as_float(b.x);
    L.smx[0][T + tid] = __int_as_float(b.y);
    L.smx[1][T + tid] = __int_as_float(b.z);
    L.smx[2][T + tid] = __int_as_float(b.w);
    L.dm[0][tid + 1] = (signed char)g.y;  // delta at the gap after the item
    if (tid == 0) L.dm[0][0] = (signed char)g.x;  // and before the first one
  }
  __syncthreads();
  // Table and tree levels for m items: lifting over levels 0..L reaches
  // 2^(L+1) - 1 positions (>= m - 1 needed), and a range query reads tree
  // nodes no larger than its m leaves.  m <= 1024 keeps L <= kItemLog.
  const int levels = min(32 - __clz((uint32_t)m), kItemLog);
  for (int lv = 1; lv <= levels; ++lv) {
    const int w = 1 << lv;
    // range minima over positions [p, p + w) for p + w <= m + 1
    if (tid + w <= m + 1)
      L.dm[lv][tid] = min(L.dm[lv - 1][tid], L.dm[lv - 1][tid + (w >> 1)]);
    if (tid < (T >> lv)) {
      const int i = (T >> lv) + tid;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        L.smn[c][i] = qmin(L.smn[c][2 * i], L.smn[c][2 * i + 1]);
        L.smx[c][i] = qmax(L.smx[c][2 * i], L.smx[c][2 * i + 1]);
      }
    }
    __syncthreads();
  }

  // 2. gap tid: its node's extent, and the node when complete
  if (tid < m - 1) {
    const int p = tid + 1;
    const int d = L.dm[0][p];
    // binary lifting on both sides at once (two independent LDS chains):
    // [cl, p) and (p, cr) hold no delta below d
    int cl = p, cr = p + 1;
    for (int lv = levels; lv >= 0; --lv) {
      const int w = 1 << lv;
      const bool gl = cl - w >= 0 && L.dm[lv][cl - w] >= d;
      const bool gr = cr + w - 1 <= m && L.dm[lv][cr] >= d;
      if (gl) cl -= w;
      if (gr) cr += w;
    }
    const int qL = cl - 1;  // -1: none inside the block
    const int qR = cr <= m ? cr : -1;
    L.qL[tid] = (short)qL;
    L.qR[tid] = (short)qR;
    if (qL >= 0 && qR >= 0) {
      L.complete[tid] = 1;
      const int a = qL, b = qR - 1;  // the node's first and last item
      float lmn[3], lmx[3], rmn[3], rmx[3];
      item_range_box(L, a, tid, lmn, lmx);
      item_range_box(L, tid + 1, b, rmn, rmx);
      const int Lr = L.il[a], s = L.ir[tid], Rr = L.ir[b];
      const int dl = L.dm[0][qL], dr = L.dm[0][qR];
      const bool root = (Lr == 0 && Rr == (int)N - 1);
      const int K = root ? 0 : (dr > dl ? Rr : Lr);
      const int lid = (Lr == s) ? -(s + 1) : s;
      const int rid = (Rr == s + 1) ? -(s + 2) : s + 1;
      int4 *node = nodes + 4 * (size_t)K;
      node[0] = pack3(lmn, lid);
      node[1] = pack3(lmx, 2);
      node[2] = pack3(rmn, rid);
      node[3] = pack3(rmx, 0);
    }
  }
  __syncthreads();

  // 3. roots of the forest, marked at their first item
  if (tid < m - 1 && L.complete[tid]) {
    const int qL = L.qL[tid], qR = L.qR[tid];
    const int a = qL, b = qR - 1;
    if (!(L.il[a] == 0 && L.ir[b] == (int)N - 1)) {  // the mesh root has none
      const int parent = L.dm[0][qR] > L.dm[0][qL] ? qR - 1 : qL - 1;
      if (!(parent >= 0 && parent <= m - 2 && L.complete[parent]))
        L.root_of[a] = tid;
    }
  }
  if (tid < m && !(L.il[tid] == 0 && L.ir[tid] == (int)N - 1)) {
    const int parent = L.dm[0][tid + 1] > L.dm[0][tid] ? tid : tid - 1;
    if (!(parent >= 0 && parent <= m - 2 && L.complete[parent]))
      L.root_of[tid] = T + tid;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kItemTile) void k_items(
    uint32_t n_all,
    const int4 *__restrict__ items_in, const uint32_t *__restrict__ count_in,
    uint32_t prev_tiles_all, int4 *items_out,  // written, then read back
    uint32_t *__restrict__ count_out, int4 *nodes_all, Ctl *ctl,
    TreeSegs segs) {
  constexpr int T = kItemTile;
  __shared__ ItemLds L;
  const int tid = threadIdx.x;
  const uint32_t gtile = xcd_tile(blockIdx.x, gridDim.x);
  uint32_t tile0 = 0, leaf_base = 0, N = n_all, prev0 = 0,
           prev_tiles = prev_tiles_all;
  if (segs.tile_seg) {
    const uint32_t sg = segs.tile_seg[gtile];
    tile0 = segs.seg_tile0[sg];
    leaf_base = segs.seg_off[sg];
    N = segs.seg_off[sg + 1] - leaf_base;
    prev0 = segs.prev_tile0[sg];
    prev_tiles = segs.prev_tiles[sg];
  }
  const uint32_t tile = gtile - tile0;  // tile within the mesh
  int4 *nodes = nodes_all + 4 * (size_t)leaf_base;
  const uint32_t fan = fan_in(prev_tiles);
  const uint32_t src0 = prev0 + tile * fan;  // first source tile
  {
    // item prefix over the source tiles: L.pre[s] = first item of tile s
    // (the total from s = nsrc on)
    const uint32_t first = tile * fan;
    const uint32_t nsrc =
        prev_tiles > first ? min(fan, prev_tiles - first) : 0u;
    const uint32_t c = (uint32_t)tid < nsrc
                           ? min(count_in[src0 + tid], (uint32_t)kFrontierMax)
                           : 0u;
    const uint32_t ex = block_exclusive_scan(c, L.wsum);
    if (tid <= kFanInMax) L.pre[tid] = ex;
  }
  __syncthreads();
  const int total = (int)L.pre[kFanInMax];
  int taken = 0;  // source items used by earlier rounds
  int carry = 0;  // items of the previous round's frontier
  while (true) {
    const int fresh = min(total - taken, T - carry);
    const bool last = taken + fresh == total;
    const int m = carry + fresh;
    item_round(L, m, carry, taken, src0, items_in, items_out, gtile, N,
               nodes);
    // the round's frontier: the next round's first items, or the output
    const int v = tid < m ? L.root_of[tid] : -1;
    const bool out = v >= 0;
    int fl = 0, fr = 0, fdl = 0, fdr = 0;
    float fmn[3] = {0.f, 0.f, 0.f}, fmx[3] = {0.f, 0.f, 0.f};
    if (out) {
      int a, b;
      if (v >= T) {
        a = b = v - T;
      } else {
        a = L.qL[v];
        b = L.qR[v] - 1;
      }
      item_range_box(L, a, b, fmn, fmx);
      fl = L.il[a];
      fr = L.ir[b];
      fdl = L.dm[0][a];
      fdr = L.dm[0][b + 1];
    }
    __syncthreads();  // every read of this round's items is done
    const uint32_t count = emit_frontier(out, fl, fr, fdl, fdr, fmn, fmx,
                                         L.wsum, gtile, items_out, count_out,
                                         ctl, tid, T);
    if (last) break;
    taken += fresh;
    carry = (int)min(count, (uint32_t)kFrontierMax);
    __syncthreads();  // the emitted items are visible to the whole block
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline uint32_t blocks_for(uint32_t n, uint32_t per) {
  uint32_t b = (n + per - 1) / per;
  return b ? b : 1;
}

static uint32_t rt_tiles(uint32_t n) {
  const uint32_t gaps = n > 1 ? n - 1 : 1;
  return blocks_for(gaps, kRtTile);
}

uint64_t radix_tree_scratch_bytes(uint32_t n) {
  const uint64_t t = rt_tiles(n);
  return t * (16ull * kRtRecords + 4 + 2 * 64 * 8);
}

void launch_radix_tree(const uint32_t *keys, uint32_t n, int32_t *parents,
                       void *scratch, hipStream_t s) {
  const uint32_t tiles = rt_tiles(n);
  int4 *recs = static_cast<int4 *>(scratch);
  unsigned long long *last =
      reinterpret_cast<unsigned long long *>(recs + (size_t)tiles * kRtRecords);
  unsigned long long *first = last + (size_t)64 * tiles;
  uint32_t *cnt = reinterpret_cast<uint32_t *>(first + (size_t)64 * tiles);
  hipLaunchKernelGGL(k_rt_tiles, dim3(tiles), dim3(kRtThreads), 0, s, keys, n,
                     parents, recs, cnt, last, first, tiles);
  if (n < 2) return;
  hipLaunchKernelGGL(k_rt_scan, dim3(128), dim3(kRtScanThreads), 0, s, last,
                     first, tiles);
  hipLaunchKernelGGL(k_rt_resolve, dim3(tiles), dim3(kRtRecords), 0, s,
                     parents, n, recs, cnt, last, first, tiles);
}

uint64_t up_pass_scratch_bytes(uint32_t n) {
  uint64_t tiles = blocks_for(n, kUpTile), items = 0, counts = 0;
  while (true) {
    items += tiles * kUpFront;
    counts += tiles;
    if (tiles == 1) break;
    tiles = (tiles + kUpFanIn - 1) / kUpFanIn;
  }
  return 48 * items + 8 * counts;  // items, item counts, completed counts
}

void launch_up_pass(const int32_t *idx, const float4 *aabb,
                    const int32_t *parents, uint32_t n, float4 *nodes,
                    void *scratch, Ctl *ctl, hipStream_t s) {
  uint32_t tiles = blocks_for(n, kUpTile);
  uint64_t total_items = 0;
  for (uint32_t t = tiles;; t = (t + kUpFanIn - 1) / kUpFanIn) {
    total_items += (uint64_t)t * kUpFront;
    if (t == 1) break;
  }
  uint64_t total_tiles = 0;
  for (uint32_t t = tiles;; t = (t + kUpFanIn - 1) / kUpFanIn) {
    total_tiles += t;
    if (t == 1) break;
  }
  int4 *items = static_cast<int4 *>(scratch);
  uint32_t *cnt = reinterpret_cast<uint32_t *>(items + 3 * total_items);
  uint32_t *done = cnt + total_tiles;
  int4 *nd = reinterpret_cast<int4 *>(nodes);
  hipLaunchKernelGGL(k_up_pass, dim3(tiles), dim3(kUpTile), 0, s, idx,
                     reinterpret_cast<const int4 *>(aabb), parents, n, nd,
                     items, cnt, done, ctl, tiles == 1 ? 1 : 0);
  while (tiles > 1) {
    const uint32_t next = (tiles + kUpFanIn - 1) / kUpFanIn;
    int4 *out = items + 3 * (size_t)tiles * kUpFront;
    hipLaunchKernelGGL(k_up_items, dim3(next), dim3(kUpTile), 0, s, items,
                       cnt, done, tiles, out, cnt + tiles, done + tiles,
                       parents, n, nd, ctl, next == 1 ? 1 : 0);
    items = out;
    cnt += tiles;
    done += tiles;
    tiles = next;
  }
}

void launch_rearrange(const float4 *tri, const int32_t *idx, uint32_t n,
                      float4 *out, hipStream_t s) {
  hipLaunchKernelGGL(k_rearrange, dim3(blocks_for(n, kMapThreads)),
                     dim3(kMapThreads), 0, s,
                     reinterpret_cast<const int4 *>(tri), idx, n,
                     reinterpret_cast<int4 *>(out));
}

// Level plan: ceil(n / tile) leaf tiles, then ceil(previous / kFanIn) item
// tiles per level, or one tile once at most kFanInMax remain (fan_in).
int tree_levels(uint32_t n, int tile, uint32_t *grid) {
  uint64_t tiles = ((uint64_t)n + tile - 1) / tile;
  int levels = 0;
  grid[levels++] = (uint32_t)(tiles ? tiles : 1);
  while (tiles > 1 && levels < kMaxTreeLevels) {
    tiles = (tiles + fan_in((uint32_t)tiles) - 1) / fan_in((uint32_t)tiles);
    grid[levels++] = (uint32_t)tiles;
  }
  return levels;
}

uint64_t tree_tiles_total(uint32_t n, int tile) {
  uint32_t grid[kMaxTreeLevels];
  const int levels = tree_levels(n, tile, grid);
  uint64_t t = 0;
  for (int L = 0; L < levels; ++L) t += grid[L];
  return t;
}

template <int T>
static void launch_leaf_level(uint32_t grid, const int4 *t4,
                              const uint32_t *skeys, const int32_t *perm,
                              uint32_t n, int4 *o4, int4 *items_out,
                              uint32_t *count_out, int4 *n4, Ctl *ctl,
                              const TreeSegs &segs, hipStream_t s) {
  hipLaunchKernelGGL(k_leaves<T>, dim3(grid), dim3(T), 0, s, t4, skeys, perm,
                     n, o4, items_out, count_out, n4, ctl, segs);
}

void launch_tree_level(int level, int tile, uint32_t grid, const float4 *tri,
                       const uint32_t *skeys, const int32_t *perm, uint32_t n,
                       float4 *otri, const int4 *items_in,
                       const uint32_t *count_in, uint32_t prev_tiles,
                       int4 *items_out, uint32_t *count_out, float4 *nodes,
                       Ctl *ctl, const TreeSegs &segs, hipStream_t s) {
  const int4 *t4 = reinterpret_cast<const int4 *>(tri);
  int4 *o4 = reinterpret_cast<int4 *>(otri);
  int4 *n4 = reinterpret_cast<int4 *>(nodes);
  if (level > 0) {
    hipLaunchKernelGGL(k_items, dim3(grid), dim3(kItemTile), 0, s, n,
                       items_in, count_in, prev_tiles, items_out, count_out,
                       n4, ctl, segs);
    return;
  }
  switch (tile) {
    case 256:
      launch_leaf_level<256>(grid, t4, skeys, perm, n, o4, items_out,
                             count_out, n4, ctl, segs, s);
      break;
    case 1024:
      launch_leaf_level<1024>(grid, t4, skeys, perm, n, o4, items_out,
                              count_out, n4, ctl, segs, s);
      break;
    default:
      launch_leaf_level<512>(grid, t4, skeys, perm, n, o4, items_out,
                             count_out, n4, ctl, segs, s);
      break;
  }
}

}  // namespace lbvh
