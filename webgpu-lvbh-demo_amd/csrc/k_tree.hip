// k_tree.hip -- radix tree, bottom-up node boxes, triangle rearrange.
//
// Reference: scripts/bvh/kernels/radix-tree.js (WGSL :137-238),
// bvh-up-pass.js (WGSL :200-280) and rearrange.js (WGSL :162-169).
//
// Two code paths produce the same BVH bytes:
//   * per-stage kernels mirroring the reference's execute() shims
//     (k_radix_tree -> parents array, k_up_pass over parents, k_rearrange);
//   * k_tree_fused, used by lbvh_build / lbvh_refit: one kernel that gathers
//     the triangles into leaf order, computes leaf boxes, and builds the
//     Karras tree bottom-up by agglomeration (Apetrei 2014) without a
//     parents array, merging inside LDS whenever both children of a split
//     lie in the block's tile and exchanging through HBM only at tile seams.
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
// Per-stage radix tree (compute_radix_tree_pointers, radix-tree.js:137-157)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int delta_at(const uint32_t *__restrict__ keys,
                                        int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = keys[i], b = keys[j];
  return (a != b) ? (int)__clz(a ^ b)
                  : 32 + (int)__clz((uint32_t)i ^ (uint32_t)j);
}

// compute_child_index (radix-tree.js:188-238) with count-leading-zeros for
// dif() (:160-169) and the textbook split search t = ceil(l/2^k) ... 1 (the
// reference's loop at :216-224 repeats t = 1 until its i32 step wraps;
// those probes do not change s).
__global__ __launch_bounds__(kMapThreads) void k_radix_tree(
    const uint32_t *__restrict__ keys, uint32_t n, int32_t *__restrict__ par) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    par[0] = 0;                       // root: never written (zero buffer)
    par[n - 1] = 0;                   // no internal node N-1
    if (n == 1) par[1] = 0;
  }
  if (i + 1 >= (int64_t)n) return;
  const int64_t N = n;
  const int d = (delta_at(keys, N, i, i + 1) - delta_at(keys, N, i, i - 1)) > 0
                    ? 1 : -1;
  const int dmin = delta_at(keys, N, i, i - d);
  int64_t lmax = 2;
  while (delta_at(keys, N, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta_at(keys, N, i, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta_at(keys, N, i, j);
  int64_t s = 0;
  for (int64_t v = 2;; v *= 2) {
    const int64_t t = (l + v - 1) / v;
    if (delta_at(keys, N, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
  const int64_t lo = (i < j) ? i : j, hi = (i < j) ? j : i;
  par[(lo == gamma) ? N + gamma : gamma] = (int32_t)i;
  par[(hi == gamma + 1) ? N + gamma + 1 : gamma + 1] = (int32_t)i;
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

// Exchanges this side of split s through HBM.  Each side record is seven
// 64-bit words (box min.xyz, max.xyz, far end), each carrying the operation's
// epoch in its high half and written with a relaxed agent-scope atomic store.
// The arrival tag is swapped relaxed too, so no cache-wide release/acquire
// fence is issued: the second arriver instead polls each sibling word until
// it shows the current epoch.  The sibling stored those words before its own
// swap, which this thread has already observed, so every poll ends; the cap
// only bounds the loop should that ever break (ctl->fault is then set).
// Returns true, with the sibling's box and far end, for the second arriver.
__device__ __forceinline__ bool arrive_global(GSide *gside, uint32_t *gtag,
                                              Ctl *ctl, uint32_t epoch,
                                              int64_t s, int side,
                                              int far_self, const float mn[3],
                                              const float mx[3], float smn[3],
                                              float smx[3], int &sfar) {
  uint64_t *mine = gside[2 * s + side].w;
  const uint64_t stamp = (uint64_t)epoch << 32;
  const uint32_t v[7] = {__float_as_uint(mn[0]), __float_as_uint(mn[1]),
                         __float_as_uint(mn[2]), __float_as_uint(mx[0]),
                         __float_as_uint(mx[1]), __float_as_uint(mx[2]),
                         (uint32_t)far_self};
#pragma unroll
  for (int k = 0; k < 7; ++k)
    __hip_atomic_store(mine + k, stamp | v[k], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  // Stores issued and acknowledged before the swap (no cache maintenance at
  // workgroup scope), so the sibling normally finds them on its first poll.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  const uint32_t old = __hip_atomic_exchange(gtag + s, epoch, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
  if (old != epoch) return false;
  const uint64_t *other = gside[2 * s + (1 - side)].w;
  uint32_t o[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    uint64_t w;
    uint32_t spins = 0;
    while (true) {
      w = __hip_atomic_load(other + k, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      if ((uint32_t)(w >> 32) == epoch) break;
      if (++spins > kSpinCap) {
        atomicOr(&ctl->fault, 2u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    o[k] = (uint32_t)w;
  }
  smn[0] = __uint_as_float(o[0]); smn[1] = __uint_as_float(o[1]);
  smn[2] = __uint_as_float(o[2]); smx[0] = __uint_as_float(o[3]);
  smx[1] = __uint_as_float(o[4]); smx[2] = __uint_as_float(o[5]);
  sfar = (int)o[6];
  return true;
}

// ---------------------------------------------------------------------------
// Per-stage up pass (bvh_upward_pass, bvh-up-pass.js:200-280)
// ---------------------------------------------------------------------------

// One thread per leaf climbs through the parents array.  Each child writes
// its half of the parent node (Karras order: a child is the left one when
// the child numbered one above it has the same parent) and publishes its box
// through the same epoch-stamped exchange as the fused kernel (arrive_global,
// keyed by the parent's id); the first arriver stops, the second merges and
// climbs.  Unlike the reference there is no zero-sentinel spin
// (bvh-up-pass.js:265-268): the stamps say when the sibling's box is there.
__global__ __launch_bounds__(kMapThreads) void k_up_pass(
    const int32_t *__restrict__ idx, const int4 *__restrict__ aabb,
    const int32_t *__restrict__ par, uint32_t n, int4 *nodes, GSide *gside,
    uint32_t *tags, Ctl *ctl) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const uint32_t epoch = ctl->epoch;
  const uint32_t src = (uint32_t)idx[g] < n ? (uint32_t)idx[g] : 0u;
  const int4 lo = aabb[2 * (size_t)src], hi = aabb[2 * (size_t)src + 1];
  float mn[3] = {__int_as_float(lo.x), __int_as_float(lo.y), __int_as_float(lo.z)};
  float mx[3] = {__int_as_float(hi.x), __int_as_float(hi.y), __int_as_float(hi.z)};
  perturb(mn, mx);
  if (n == 1) {
    write_single_leaf_node(nodes, mn, mx);
    return;
  }
  if (g == n - 1) {
    const int4 z = make_int4(0, 0, 0, 0);
    int4 *last = nodes + 4 * (size_t)(n - 1);
    last[0] = z; last[1] = z; last[2] = z; last[3] = z;
  }
  bool leaf = true;
  int64_t num = g;  // leaf position or internal node id
  int32_t p = par[n + g];
  // A valid tree is at most 64 levels deep (each level consumes a bit of
  // the 62-bit position-augmented key); the step cap and the range check
  // keep a malformed parents array from running away.
  for (int step = 0; step < 128; ++step) {
    if (p < 0 || p >= (int32_t)n - 1) return;
    const int64_t sib = num + 1;
    const bool left = (sib <= (int64_t)n - 1 && par[n + sib] == p) ||
                      (sib <= (int64_t)n - 2 && par[sib] == p);
    const int id = leaf ? -(int)(num + 1) : (int)num;
    int4 *node = nodes + 4 * (size_t)p;
    if (left) {
      node[0] = pack3(mn, id);
      node[1] = pack3(mx, 2);
    } else {
      node[2] = pack3(mn, id);
      node[3] = pack3(mx, 0);
    }
    float smn[3], smx[3];
    int unused;
    if (!arrive_global(gside, tags, ctl, epoch, p, left ? 0 : 1, 0, mn, mx,
                       smn, smx, unused))
      return;  // first arriver
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], smn[k]);
      mx[k] = fmaxf(mx[k], smx[k]);
    }
    if (p == 0) return;  // root merged
    leaf = false;
    num = p;
    p = par[p];
  }
}

// rearrange_triangles (rearrange.js:162-169): whole 48-byte records.
__global__ __launch_bounds__(kMapThreads) void k_rearrange(
    const int4 *__restrict__ tri, const int32_t *__restrict__ idx, uint32_t n,
    int4 *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = (uint32_t)idx[i];
  const size_t s = 3 * (size_t)(j < n ? j : 0u);
  const int4 a = tri[s], b = tri[s + 1], c = tri[s + 2];
  out[3 * (size_t)i] = a;
  out[3 * (size_t)i + 1] = b;
  out[3 * (size_t)i + 2] = c;
}

// ---------------------------------------------------------------------------
// Fused leaf + tree + up pass
// ---------------------------------------------------------------------------

__device__ __forceinline__ int delta_next(uint32_t a, uint32_t b, uint32_t i) {
  return (a != b) ? (int)__clz(a ^ b) : 32 + (int)__clz(i ^ (i + 1u));
}

template <int T>
struct TreeLds {
  int delta[T + 1];         // delta(ts-1+k, ts+k), k = 0..T
  float box[2][6][T];       // [side][min.xyz max.xyz][split slot]
  int far[2][T];            // far end of the side's leaf range
  uint32_t cnt[T];          // local arrivals per split slot
};

// State of one climbing thread: the completed node [l, r] and its box, plus
// the halves of that node once its children merged (written when its id is
// known, i.e. once its parent's side is decided).
struct Climb {
  int64_t l, r;
  float mn[3], mx[3];
  bool have_node;
  int4 h[4];
  bool global;
  // pending local arrival
  bool pend;
  int64_t ps;
  int pside;
};

// Climbs until the thread stops: first arrival at a split (locally pending
// or globally done) or the root.  `resume` re-enters a pending local
// arrival through HBM.
template <int T>
__device__ void climb(Climb &c, TreeLds<T> &L,
                      const uint32_t *__restrict__ skeys, int64_t n,
                      int64_t ts, int64_t te, int4 *nodes, GSide *gside,
                      uint32_t *gtag, Ctl *ctl, uint32_t epoch, bool resume) {
  while (true) {
    int64_t s;
    int side, far_self;
    if (resume) {
      resume = false;
      s = c.ps;
      side = c.pside;
      far_self = side == 0 ? (int)c.l : (int)c.r;
    } else {
      auto delta = [&](int64_t x) -> int {  // x in [0, n-2]
        if (x >= ts - 1 && x <= te - 1) return L.delta[x - ts + 1];
        return delta_next(skeys[x], skeys[x + 1], (uint32_t)x);
      };
      const int dl = (c.l == 0) ? -1 : delta(c.l - 1);
      const int dr = (c.r == n - 1) ? -1 : delta(c.r);
      const bool root = (c.l == 0 && c.r == n - 1);
      const bool is_left = dr > dl;
      if (c.have_node) {
        const int64_t K = root ? 0 : (is_left ? c.r : c.l);
        int4 *node = nodes + 4 * K;
        node[0] = c.h[0]; node[1] = c.h[1]; node[2] = c.h[2]; node[3] = c.h[3];
        c.have_node = false;
      }
      if (root) return;
      s = is_left ? c.r : c.l - 1;
      side = is_left ? 0 : 1;
      far_self = is_left ? (int)c.l : (int)c.r;
    }
    float smn[3], smx[3];
    int sfar;
    if (!c.global && s >= ts && s + 2 <= te) {
      const int64_t slot = s - ts;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        L.box[side][k][slot] = c.mn[k];
        L.box[side][3 + k][slot] = c.mx[k];
      }
      L.far[side][slot] = far_self;
      const uint32_t old = __hip_atomic_fetch_add(
          &L.cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) {
        c.pend = true;
        c.ps = s;
        c.pside = side;
        return;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        smn[k] = L.box[1 - side][k][slot];
        smx[k] = L.box[1 - side][3 + k][slot];
      }
      sfar = L.far[1 - side][slot];
    } else {
      c.global = true;
      if (!arrive_global(gside, gtag, ctl, epoch, s, side, far_self, c.mn,
                         c.mx, smn, smx, sfar))
        return;
    }
    // Merge into the split-s node: children [L, s] and [s+1, R].
    const bool self_left = (side == 0);
    const int64_t Lr = self_left ? c.l : (int64_t)sfar;
    const int64_t Rr = self_left ? (int64_t)sfar : c.r;
    const float *lmn = self_left ? c.mn : smn, *lmx = self_left ? c.mx : smx;
    const float *rmn = self_left ? smn : c.mn, *rmx = self_left ? smx : c.mx;
    const int lid = (Lr == s) ? -(int)(s + 1) : (int)s;
    const int rid = (Rr == s + 1) ? -(int)(s + 2) : (int)(s + 1);
    c.h[0] = pack3(lmn, lid);
    c.h[1] = pack3(lmx, 2);
    c.h[2] = pack3(rmn, rid);
    c.h[3] = pack3(rmx, 0);
    c.have_node = true;
    float nmn[3], nmx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      nmn[k] = fminf(lmn[k], rmn[k]);
      nmx[k] = fmaxf(lmx[k], rmx[k]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { c.mn[k] = nmn[k]; c.mx[k] = nmx[k]; }
    c.l = Lr;
    c.r = Rr;
  }
}

// One block per tile of T sorted positions.  Phase 1: every thread gathers
// its triangle into leaf order (O_TRI), takes its leaf box and climbs,
// merging through LDS while both children of a split lie in the tile; a
// split at a tile seam goes through HBM at once.  Phase 2 (after the block
// barrier): a first arriver whose local sibling never came (the sibling
// subtree reaches outside the tile) re-arrives through HBM.  Above the seams
// every step is a global exchange.  No thread waits for another thread's
// progress, only for stores already issued.
template <int T>
__global__ __launch_bounds__(T) void k_tree_fused(
    const int4 *__restrict__ tri, const uint32_t *__restrict__ skeys,
    const int32_t *__restrict__ perm, uint32_t n, int4 *__restrict__ otri,
    int4 *nodes, GSide *gside, uint32_t *gtag, Ctl *ctl) {
  __shared__ TreeLds<T> L;
  const int64_t t = threadIdx.x;
  const int64_t ts = (int64_t)blockIdx.x * T;
  const int64_t i = ts + t;
  const int64_t N = n;
  const int64_t te = (ts + T < N) ? ts + T : N;
  const uint32_t epoch = ctl->epoch;

  L.cnt[t] = 0u;
  Climb c;
  c.have_node = false;
  c.global = false;
  c.pend = false;
  if (i < N) {
    const uint32_t k0 = skeys[i];
    L.delta[t + 1] = (i + 1 < N) ? delta_next(k0, skeys[i + 1], (uint32_t)i) : -1;
    if (t == 0)
      L.delta[0] = (ts == 0) ? -1
                             : delta_next(skeys[ts - 1], k0, (uint32_t)(ts - 1));
    const uint32_t pi = (uint32_t)perm[i];
    const size_t src = 3 * (size_t)(pi < (uint32_t)N ? pi : 0u);
    const int4 a = tri[src], b = tri[src + 1], d = tri[src + 2];
    otri[3 * i] = a;
    otri[3 * i + 1] = b;
    otri[3 * i + 2] = d;
    tri_box(a, b, d, c.mn, c.mx);
    perturb(c.mn, c.mx);
    c.l = c.r = i;
  }
  __syncthreads();
  if (N == 1) {
    if (i == 0) write_single_leaf_node(nodes, c.mn, c.mx);
    return;
  }
  if (i == N - 1) {
    const int4 z = make_int4(0, 0, 0, 0);
    int4 *last = nodes + 4 * (N - 1);
    last[0] = z; last[1] = z; last[2] = z; last[3] = z;
  }
  if (i < N) climb<T>(c, L, skeys, N, ts, te, nodes, gside, gtag, ctl, epoch, false);
  __syncthreads();
  if (c.pend && L.cnt[c.ps - ts] == 1u) {
    c.pend = false;
    c.global = true;
    climb<T>(c, L, skeys, N, ts, te, nodes, gside, gtag, ctl, epoch, true);
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline uint32_t blocks_for(uint32_t n, uint32_t per) {
  uint32_t b = (n + per - 1) / per;
  return b ? b : 1;
}

void launch_radix_tree(const uint32_t *keys, uint32_t n, int32_t *parents,
                       hipStream_t s) {
  const uint32_t threads = n > 1 ? n - 1 : 1;
  hipLaunchKernelGGL(k_radix_tree, dim3(blocks_for(threads, kMapThreads)),
                     dim3(kMapThreads), 0, s, keys, n, parents);
}

void launch_up_pass(const int32_t *idx, const float4 *aabb,
                    const int32_t *parents, uint32_t n, float4 *nodes,
                    GSide *gside, uint32_t *tags, Ctl *ctl, hipStream_t s) {
  hipLaunchKernelGGL(k_up_pass, dim3(blocks_for(n, kMapThreads)),
                     dim3(kMapThreads), 0, s, idx,
                     reinterpret_cast<const int4 *>(aabb), parents, n,
                     reinterpret_cast<int4 *>(nodes), gside, tags, ctl);
}

void launch_rearrange(const float4 *tri, const int32_t *idx, uint32_t n,
                      float4 *out, hipStream_t s) {
  hipLaunchKernelGGL(k_rearrange, dim3(blocks_for(n, kMapThreads)),
                     dim3(kMapThreads), 0, s,
                     reinterpret_cast<const int4 *>(tri), idx, n,
                     reinterpret_cast<int4 *>(out));
}

void launch_tree_fused(const float4 *tri, const uint32_t *skeys,
                       const int32_t *perm, uint32_t n, float4 *otri,
                       float4 *nodes, GSide *gside, uint32_t *gtag, Ctl *ctl,
                       int tile, hipStream_t s) {
  const int4 *t4 = reinterpret_cast<const int4 *>(tri);
  int4 *o4 = reinterpret_cast<int4 *>(otri);
  int4 *n4 = reinterpret_cast<int4 *>(nodes);
  switch (tile) {
    case 256:
      hipLaunchKernelGGL(k_tree_fused<256>, dim3(blocks_for(n, 256)),
                         dim3(256), 0, s, t4, skeys, perm, n, o4, n4, gside,
                         gtag, ctl);
      break;
    case 1024:
      hipLaunchKernelGGL(k_tree_fused<1024>, dim3(blocks_for(n, 1024)),
                         dim3(1024), 0, s, t4, skeys, perm, n, o4, n4, gside,
                         gtag, ctl);
      break;
    default:
      hipLaunchKernelGGL(k_tree_fused<512>, dim3(blocks_for(n, 512)),
                         dim3(512), 0, s, t4, skeys, perm, n, o4, n4, gside,
                         gtag, ctl);
      break;
  }
}

}  // namespace lbvh
