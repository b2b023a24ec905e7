// lbvh_api.hip -- host side of the C ABI declared in include/lbvh.h.
//
// A context owns one HIP stream (default priority), a device control block
// (Ctl) and a grow-only scratch arena.  All device memory it hands out or
// keeps is charged against the context's HBM budget.  Entry points validate
// their arguments on the host, enqueue kernels and return; nothing here
// blocks on the GPU except lbvh_upload/lbvh_download/lbvh_synchronize.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lbvh.h"
#include "alloc_table.h"
#include "lbvh_internal.h"

namespace lbvh {
void launch_synth_uniform(uint64_t seed, uint32_t n, double half_extent,
                          float *tri, hipStream_t s);
}

using namespace lbvh;

namespace {

struct DevBuf {
  void *ptr = nullptr;
  uint64_t bytes = 0;
};

}  // namespace

struct lbvh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;  // 0 = unlimited
  uint64_t in_use = 0;
  int div_mode = LBVH_DIV_MUL_RCP;
  int tree_tile = kTreeTileDefault;
  bool tree_tile_fixed = false;  // set_tuning overrides the tuned table
  int tune_tile[33] = {0};       // autotuned tree tile per log2 size bucket
  int sort_bits = kRadixBits;    // sort digit width (8 or 10)
  bool sort_bits_fixed = false;
  int tune_bits[33] = {0};       // autotuned digit width per size bucket
  std::string last_error;
  AllocTable user_allocs;  // buffers handed out by lbvh_alloc
  std::recursive_mutex mu;
  // device state
  Ctl *ctl = nullptr;
  DevBuf keys_a, keys_b, idx_a, idx_b;  // sort ping-pong, Morton output
  DevBuf digits;                        // sort tile counts + offsets
  DevBuf up_items;                      // per-stage up-pass items and counts
  DevBuf rt_scratch;                    // per-stage radix tree records, summaries
  DevBuf tree_items, tree_counts;       // agglomeration item slots, counts
  DevBuf partials;                      // bounds partials
  DevBuf cen;                           // triangle box centres (GPU bounds)
  // batched builds: scratch, and the plan (tile tables + captured graph) of
  // the last batch layout
  DevBuf b_keys_a, b_keys_b, b_idx_a, b_idx_b, b_digits, b_ord,
      b_bounds, b_tree_items, b_tree_counts, b_table, b_cen;
  struct BatchPlan {
    std::vector<uint32_t> offsets;
    int tile = 0, bits = kRadixBits;
    uint32_t S = 0, total = 0, sort_tiles = 0;
    int levels = 0;
    std::vector<uint32_t> level_tiles;
    // u32 offsets of the tables inside b_table
    uint64_t o_seg_off = 0, o_tile_seg = 0, o_seg_tile0 = 0;
    std::vector<uint64_t> o_ltile_seg, o_lseg_tile0, o_lprev0, o_lprev_n;
    uint64_t tree_tiles = 0;  // all levels
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> graph_key;
  } plan;
  // profiling: event pairs bracketing each stage, read back on demand
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  struct Mark {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  lbvh_profile acc{};
  // pinned staging buffers for large host copies, two per copy thread,
  // allocated on the first large copy
  std::vector<void *> stage_buf;
  std::vector<hipEvent_t> stage_ev;
  // device address of this device's static scratch (lbvh_build_static), or
  // null if the lookup failed
  void *static_scratch = nullptr;
};

namespace {

// Host copies of at least kStageMin bytes go through pinned staging: a
// pageable hipMemcpy copies through the runtime's own staging on one thread
// (about 30 GB/s host-to-device measured on MI355X), where kStageThreads
// threads filling pinned kStageChunk buffers keep the DMA engine busy
// (about 50 GB/s; pinned DMA alone peaks near 58 GB/s).
constexpr uint64_t kStageMin = 64ull << 20;
constexpr uint64_t kStageChunk = 16ull << 20;
constexpr int kStageThreads = 4;

// Scratch of lbvh_build_static: one array per device, in the code object,
// allocated when the library's code is loaded on the device and held until
// the process ends.  Host code never passes `scratch` itself (a host-side
// stand-in) to a kernel: it looks up the array's device address once per
// device (static_scratch_address) and uses that.
constexpr uint32_t kStaticMaxTris = 1u << 20;
constexpr uint64_t kStaticScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
void *g_static_ptr[kMaxDevices] = {};
bool g_static_looked_up[kMaxDevices] = {};

// The calling thread's current device must be `device`.
void *static_scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_static_mu);
  if (!g_static_looked_up[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) {
      g_static_ptr[device] = p;
    } else {
      (void)hipGetLastError();
    }
    g_static_looked_up[device] = true;
  }
  return g_static_ptr[device];
}

int fail(lbvh_ctx *ctx, int status, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return status;
}

// Allocates the staging buffers once; false (and nothing held) when pinned
// memory is not available, so the caller copies directly instead.
bool ensure_staging(lbvh_ctx *ctx) {
  if (!ctx->stage_buf.empty()) return true;
  std::vector<void *> bufs;
  std::vector<hipEvent_t> evs;
  bool ok = true;
  for (int i = 0; ok && i < 2 * kStageThreads; ++i) {
    void *p = nullptr;
    hipEvent_t e = nullptr;
    ok = hipHostMalloc(&p, kStageChunk, 0) == hipSuccess;
    if (ok) bufs.push_back(p);
    ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (ok) evs.push_back(e);
  }
  if (!ok) {
    for (void *p : bufs) (void)hipHostFree(p);
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    (void)hipGetLastError();
    return false;
  }
  ctx->stage_buf = std::move(bufs);
  ctx->stage_ev = std::move(evs);
  return true;
}

void release_staging(lbvh_ctx *ctx) {
  for (void *p : ctx->stage_buf) (void)hipHostFree(p);
  for (hipEvent_t e : ctx->stage_ev) (void)hipEventDestroy(e);
  ctx->stage_buf.clear();
  ctx->stage_ev.clear();
}

// A large host <-> device copy through the staging buffers on the context's
// stream, synchronous.  Copy thread i takes chunks i, i + kStageThreads, ...
// through its buffers 2i and 2i + 1: to the device it fills a buffer (after
// the DMA that last read it is done) and queues its DMA; to the host it
// queues the DMA into a buffer and copies the previous chunk out meanwhile.
hipError_t staged_copy(lbvh_ctx *ctx, char *dst, const char *src,
                       uint64_t bytes, bool to_device) {
  const uint64_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
  std::atomic<int> err{(int)hipSuccess};
  const auto note = [&](hipError_t e) {
    int ok = (int)hipSuccess;
    if (e != hipSuccess) err.compare_exchange_strong(ok, (int)e);
    return e == hipSuccess;
  };
  const auto worker = [&](int i) {
    if (!note(hipSetDevice(ctx->device))) return;
    uint64_t prev = UINT64_MAX;  // to the host: the chunk still in a buffer
    int prev_b = 0;
    int j = 0;
    for (uint64_t c = i; c < chunks; c += kStageThreads, ++j) {
      const int b = 2 * i + (j & 1);
      const uint64_t off = c * kStageChunk;
      const uint64_t len = std::min<uint64_t>(kStageChunk, bytes - off);
      char *buf = static_cast<char *>(ctx->stage_buf[b]);
      if (j >= 2 && !note(hipEventSynchronize(ctx->stage_ev[b]))) return;
      if (to_device) {
        memcpy(buf, src + off, len);
        if (!note(hipMemcpyAsync(dst + off, buf, len, hipMemcpyHostToDevice,
                                 ctx->stream)))
          return;
      } else if (!note(hipMemcpyAsync(buf, src + off, len,
                                      hipMemcpyDeviceToHost, ctx->stream))) {
        return;
      }
      if (!note(hipEventRecord(ctx->stage_ev[b], ctx->stream))) return;
      if (!to_device && prev != UINT64_MAX) {
        if (!note(hipEventSynchronize(ctx->stage_ev[prev_b]))) return;
        const uint64_t poff = prev * kStageChunk;
        memcpy(dst + poff, ctx->stage_buf[prev_b],
               std::min<uint64_t>(kStageChunk, bytes - poff));
      }
      prev = c;
      prev_b = b;
    }
    if (!to_device && prev != UINT64_MAX) {
      if (!note(hipEventSynchronize(ctx->stage_ev[prev_b]))) return;
      const uint64_t poff = prev * kStageChunk;
      memcpy(dst + poff, ctx->stage_buf[prev_b],
             std::min<uint64_t>(kStageChunk, bytes - poff));
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(kStageThreads);
  for (int i = 0; i < kStageThreads; ++i) pool.emplace_back(worker, i);
  for (auto &t : pool) t.join();
  const hipError_t s = hipStreamSynchronize(ctx->stream);
  return err.load() != (int)hipSuccess ? (hipError_t)err.load() : s;
}

int hip_fail(lbvh_ctx *ctx, hipError_t e, const char *what) {
  return fail(ctx, LBVH_ERR_DEVICE,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                          \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

bool fits_budget(const lbvh_ctx *ctx, uint64_t extra) {
  return ctx->budget == 0 || ctx->in_use + extra <= ctx->budget;
}

int dev_alloc(lbvh_ctx *ctx, uint64_t bytes, void **out) {
  if (!fits_budget(ctx, bytes))
    return fail(ctx, LBVH_ERR_OUT_OF_BUDGET,
                "allocation of " + std::to_string(bytes) +
                    " bytes exceeds the HBM budget (" +
                    std::to_string(ctx->in_use) + " of " +
                    std::to_string(ctx->budget) + " in use)");
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, LBVH_ERR_OUT_OF_MEMORY,
                "hipMalloc(" + std::to_string(bytes) + ") failed: " +
                    hipGetErrorString(e));
  }
  ctx->in_use += bytes;
  *out = p;
  return LBVH_OK;
}

void dev_free_raw(lbvh_ctx *ctx, void *p, uint64_t bytes) {
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(p);
  ctx->in_use -= bytes;
}

void dev_release(lbvh_ctx *ctx, DevBuf &b) {
  if (b.ptr) {
    (void)hipFree(b.ptr);
    ctx->in_use -= b.bytes;
  }
  b.ptr = nullptr;
  b.bytes = 0;
}

// Grow-only scratch buffer (zeroed when (re)allocated if `zero`).
int ensure(lbvh_ctx *ctx, DevBuf &b, uint64_t bytes, bool zero) {
  if (b.bytes >= bytes) return LBVH_OK;
  // Grow with headroom so that slowly increasing sizes do not reallocate on
  // every call.
  uint64_t want = bytes + bytes / 4;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  dev_release(ctx, b);
  void *p = nullptr;
  int st = dev_alloc(ctx, want, &p);
  if (st == LBVH_ERR_OUT_OF_BUDGET && want > bytes) {
    want = bytes;
    st = dev_alloc(ctx, want, &p);
  }
  if (st != LBVH_OK) return st;
  b.ptr = p;
  b.bytes = want;
  if (zero) HIP_TRY(ctx, hipMemsetAsync(p, 0, want, ctx->stream));
  return LBVH_OK;
}

// Stage bracketing for lbvh_profile_*: a StageScope records one event before
// the stage's launches and one after them when profiling is enabled.
hipEvent_t take_event(lbvh_ctx *ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct StageScope {
  lbvh_ctx *ctx;
  int stage;
  hipEvent_t a = nullptr;
  StageScope(lbvh_ctx *c, int s) : ctx(c), stage(s) {
    if (ctx->prof && (a = take_event(ctx)))
      (void)hipEventRecord(a, ctx->stream);
  }
  ~StageScope() {
    if (!a) return;
    hipEvent_t b = take_event(ctx);
    if (!b) {
      ctx->ev_pool.push_back(a);
      return;
    }
    (void)hipEventRecord(b, ctx->stream);
    ctx->marks.push_back({stage, a, b});
  }
};

std::vector<DevBuf *> all_scratch(lbvh_ctx *ctx) {
  return {&ctx->keys_a,  &ctx->keys_b,   &ctx->idx_a,    &ctx->idx_b,
          &ctx->digits,  &ctx->up_items,    &ctx->rt_scratch, &ctx->tree_items,
          &ctx->tree_counts, &ctx->partials, &ctx->cen,
          &ctx->b_keys_a,
          &ctx->b_keys_b, &ctx->b_idx_a, &ctx->b_idx_b,  &ctx->b_digits,
          &ctx->b_ord,    &ctx->b_bounds, &ctx->b_tree_items,
          &ctx->b_tree_counts, &ctx->b_table, &ctx->b_cen};
}

void drop_plan(lbvh_ctx *ctx) {
  if (ctx->plan.exec) (void)hipGraphExecDestroy(ctx->plan.exec);
  if (ctx->plan.graph) (void)hipGraphDestroy(ctx->plan.graph);
  ctx->plan = lbvh_ctx::BatchPlan{};
}

int check_launch(lbvh_ctx *ctx, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return LBVH_OK;
}

int check_n(lbvh_ctx *ctx, uint32_t n) {
  if (n > LBVH_MAX_TRIS)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT,
                "triangle count " + std::to_string(n) + " exceeds the maximum " +
                    std::to_string(LBVH_MAX_TRIS));
  return LBVH_OK;
}

// A device pointer argument and the bytes an entry point reads or writes
// from it.  Pointers inside an lbvh_alloc buffer must have room for those
// bytes; other pointers (the caller's own allocations) are not checked.
struct Span {
  const void *p;
  uint64_t bytes;
  const char *what;
};
int check_spans(lbvh_ctx *ctx, std::initializer_list<Span> spans) {
  for (const Span &s : spans) {
    uint64_t room = 0;
    if (s.p && s.bytes &&
        ctx->user_allocs.check(s.p, s.bytes, &room) == AllocTable::kOverrun)
      return fail(ctx, LBVH_ERR_SIZE_MISMATCH,
                  std::string(s.what) + " needs " + std::to_string(s.bytes) +
                      " bytes, but its lbvh_alloc buffer holds " +
                      std::to_string(room) + " from that address");
  }
  return LBVH_OK;
}

// Caller bounds (MODEL_BOUNDS, aabb-z-idx.js:113-124) must be finite: the
// reference would quantise against NaN or infinite bounds into keys WGSL
// leaves implementation-defined.  min == max on an axis is accepted and
// deterministic: that axis's coordinate is 0 for centres on the plane (0/0
// converts to 0) and saturates off it.
int check_bounds(lbvh_ctx *ctx, const float *bounds, uint32_t count) {
  if (!bounds) return LBVH_OK;
  for (uint64_t i = 0; i < 6ull * count; ++i)
    if (!std::isfinite(bounds[i]))
      return fail(ctx, LBVH_ERR_INVALID_ARGUMENT,
                  "caller bounds must be finite (entry " + std::to_string(i) +
                      " is " + std::to_string(bounds[i]) + ")");
  return LBVH_OK;
}

int ensure_sort_scratch(lbvh_ctx *ctx, uint32_t n, int bits) {
  int st;
  if ((st = ensure(ctx, ctx->keys_a, 4ull * n, false))) return st;
  if ((st = ensure(ctx, ctx->keys_b, 4ull * n, false))) return st;
  if ((st = ensure(ctx, ctx->idx_a, 4ull * n, false))) return st;
  if ((st = ensure(ctx, ctx->idx_b, 4ull * n, false))) return st;
  // per-tile digit counts and offsets of one pass
  const uint64_t words = (uint64_t)sort_tiles(n) * sort_radix(bits) * 2;
  return ensure(ctx, ctx->digits, 4ull * words, false);
}

int bucket_of(uint32_t n) {
  int b = 0;
  while ((2ull << b) <= n && b < 32) ++b;
  return b;  // floor(log2 n)
}

// A tunable's value for an n-triangle build: a fixed value, else the
// autotuned entry of n's size bucket (or the nearest tuned bucket), else
// `untuned` (0: the context's default value).
int tuned_for(int value, bool fixed, const int (&table)[33], uint32_t n,
              int untuned = 0) {
  if (fixed || n == 0) return value;
  const int b = bucket_of(n);
  for (int d = 0; d <= 32; ++d) {
    if (b - d >= 0 && table[b - d]) return table[b - d];
    if (b + d <= 32 && table[b + d]) return table[b + d];
  }
  return untuned ? untuned : value;
}

int tile_for(const lbvh_ctx *ctx, uint32_t n) {
  return tuned_for(ctx->tree_tile, ctx->tree_tile_fixed, ctx->tune_tile, n);
}

// Untuned digit widths, from builds of uniform meshes on MI355X (best
// leaf tile each; sorts of at most 2^20 keys use 1,024-key tiles):
//   keys        8-bit     10-bit
//   2^16        0.077     0.076 ms
//   2^17        0.087     0.087
//   2^18        0.100     0.103
//   2^19        0.125     0.131
//   2^20        0.170     0.188
//   2^21        0.260     0.256   (4,096-key tiles from here on)
//   2^22        0.411     0.425
// 10-bit digits save a pass but carry four times the count words per tile,
// which small tiles multiply; so 10-bit up to 2^17 keys and from 2^20 + 1 to
// 2^21, 8-bit elsewhere.
int untuned_bits(uint32_t n) {
  const bool wide = n <= (1u << 17) || (n > (1u << 20) && n <= (1u << 21));
  return wide ? kWideBits : kRadixBits;
}

int bits_for(const lbvh_ctx *ctx, uint32_t n) {
  return tuned_for(ctx->sort_bits, ctx->sort_bits_fixed, ctx->tune_bits, n,
                   untuned_bits(n));
}

// The key buffers of each pass of a sort from `in` to `out`: each pass reads
// one buffer and writes another.  An even number of passes runs
// in -> tmp -> out -> tmp -> out (`in` may equal `out`); an odd one runs
// in -> out -> tmp -> out when in != out (tmp may equal in), else
// in -> tmp -> x -> out.  Payloads: generated -> ... ending in vout, through
// vtmp.
struct SortPlan {
  const uint32_t *src[4];
  uint32_t *dst[4];
  const int32_t *vsrc[4];
  int32_t *vdst[4];
  int passes;
};
SortPlan sort_plan(int bits, int key_bits, const uint32_t *in, uint32_t *tmp,
                   uint32_t *out, uint32_t *x, int32_t *vtmp, int32_t *vout) {
  SortPlan P{};
  P.passes = sort_passes(bits, key_bits);
  uint32_t *chain[5];
  int32_t *vchain[5];
  if (P.passes % 2 == 0) {
    uint32_t *c[5] = {const_cast<uint32_t *>(in), tmp, out, tmp, out};
    std::copy(c, c + 5, chain);
  } else if (in != out) {
    uint32_t *c[5] = {const_cast<uint32_t *>(in), out, tmp, out, nullptr};
    std::copy(c, c + 5, chain);
  } else {
    uint32_t *c[5] = {const_cast<uint32_t *>(in), tmp, x, out, nullptr};
    std::copy(c, c + 5, chain);
  }
  // payloads end in vout: alternate backwards from the last pass
  for (int p = P.passes; p >= 1; --p)
    vchain[p] = ((P.passes - p) % 2 == 0) ? vout : vtmp;
  vchain[0] = nullptr;
  for (int p = 0; p < P.passes; ++p) {
    P.src[p] = chain[p];
    P.dst[p] = chain[p + 1];
    P.vsrc[p] = vchain[p];
    P.vdst[p] = vchain[p + 1];
  }
  return P;
}

int ensure_tree_scratch(lbvh_ctx *ctx, uint32_t n, int tile) {
  int st;
  const uint64_t tiles = tree_tiles_total(n, tile);
  if ((st = ensure(ctx, ctx->tree_items, 16ull * kItemInt4 * kTreeFrontier * tiles,
                   false)))
    return st;
  return ensure(ctx, ctx->tree_counts, 4ull * tiles, false);
}

int ensure_partials(lbvh_ctx *ctx) {
  return ensure(ctx, ctx->partials, 32ull * kBoundsBlocks, false);
}

// The digit passes of one sort (sort_plan: `in` may equal `out`, as the
// stand-alone sort is in place like the reference's; x is used only then,
// with an odd pass count).  counted0: the Morton kernel already wrote pass
// 0's per-tile counts into the digit buffer; the other passes count their
// own.  key_bits: significant bits of the keys (30 for Morton keys).
void run_sort(lbvh_ctx *ctx, uint32_t *digits, const uint32_t *in,
              uint32_t *tmp, uint32_t *out, uint32_t *x, int32_t *vtmp,
              int32_t *vout, uint32_t n, bool counted0, int bits,
              int key_bits = kKeyBits) {
  const uint32_t tiles = sort_tiles(n);
  uint32_t *cnt = digits;
  uint32_t *off = cnt + (size_t)sort_radix(bits) * tiles;
  const SortPlan P = sort_plan(bits, key_bits, in, tmp, out, x, vtmp, vout);
  for (int p = 0; p < P.passes; ++p) {
    StageScope sc(ctx, LBVH_STAGE_SORT0 + p);
    launch_sort_pass(bits, p, P.src[p], P.vsrc[p], P.dst[p], P.vdst[p], n,
                     tiles, cnt, off, p == 0 && counted0, ctx->ctl,
                     ctx->stream, nullptr, 0, key_bits);
  }
}

// The tree levels of one mesh (ensure_tree_scratch done): the leaf tiles,
// then the item levels above them, each level its own profiling stage.
void run_tree(lbvh_ctx *ctx, int4 *items, uint32_t *counts, const float *tri,
              const uint32_t *skeys, const int32_t *perm, uint32_t n,
              float *tri_out, void *bvh_out, int tile) {
  uint32_t grid[kMaxTreeLevels];
  const int levels = tree_levels(n, tile, grid);
  uint64_t base = 0, prev = 0;
  for (int L = 0; L < levels; ++L) {
    StageScope sc(ctx, L ? LBVH_STAGE_TREE_TOP : LBVH_STAGE_TREE);
    launch_tree_level(L, tile, grid[L], reinterpret_cast<const float4 *>(tri),
                      skeys, perm, n, reinterpret_cast<float4 *>(tri_out),
                      items + kItemInt4 * kTreeFrontier * prev, counts + prev,
                      L ? grid[L - 1] : 0u, items + kItemInt4 * kTreeFrontier * base,
                      counts + base, static_cast<float4 *>(bvh_out), ctx->ctl,
                      TreeSegs{}, ctx->stream);
    prev = base;
    base += grid[L];
  }
}

MortonParams morton_params(const lbvh_ctx *ctx, const float *bounds) {
  MortonParams p{};
  for (int c = 0; c < 3; ++c) {
    p.bmin[c] = bounds[c];
    volatile float ext = bounds[3 + c] - bounds[c];  // one fp32 rounding
    p.ext[c] = ext;
    volatile float rcp = 1.0f / p.ext[c];
    p.rcp[c] = rcp;
  }
  p.div_mode = ctx->div_mode;
  return p;
}

#define CTX_GUARD(ctx)                                        \
  if (!(ctx)) return LBVH_ERR_INVALID_ARGUMENT;               \
  std::lock_guard<std::recursive_mutex> lock_((ctx)->mu);     \
  (void)hipSetDevice((ctx)->device)


// ---- the single-mesh build pipeline -------------------------------------------

// Device scratch of one build: sort ping-pong buffers, per-tile digit
// counts and offsets, agglomeration item slots and counts, bounds partials
// and (GPU bounds) the box centres.  From the context's arena (lbvh_build)
// or the library's static array (lbvh_build_static).
struct BuildScratch {
  uint32_t *keys_a = nullptr, *keys_b = nullptr;
  int32_t *idx_a = nullptr, *idx_b = nullptr;
  uint32_t *digits = nullptr;
  int4 *tree_items = nullptr;
  uint32_t *tree_counts = nullptr;
  float4 *partials = nullptr;
  float *cen = nullptr;
};

int check_build_args(lbvh_ctx *ctx, const float *tri, uint32_t n,
                     const void *bvh_out, const float *tri_out,
                     const uint32_t *sorted_keys_out, const int32_t *perm_out) {
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (n == 0) return LBVH_OK;
  if (!tri || !bvh_out || !tri_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  return check_spans(ctx, {{tri, 48ull * n, "triangles"},
                           {bvh_out, 64ull * n, "BVH output"},
                           {tri_out, 48ull * n, "triangle output"},
                           {sorted_keys_out, 4ull * n, "sorted-key output"},
                           {perm_out, 4ull * n, "permutation output"}});
}

// Enqueues the whole build of n >= 1 triangles: bounds (unless given),
// Morton keys with the first pass's digit counts, the sort, the
// agglomeration levels.  Enqueues kernels only: no allocation, copy or
// memset.
void build_pipeline(lbvh_ctx *ctx, const BuildScratch &S, const float *tri,
                    uint32_t n, const float *bounds_host, void *bvh_out,
                    float *tri_out, uint32_t *sorted_keys_out,
                    int32_t *perm_out, int bits, int tile) {
  const bool sort = n > 1;
  // GPU bounds: the bounds pass also writes the box centres, and the Morton
  // pass reads those (12 bytes per triangle) instead of the triangles
  const bool use_cen = S.cen != nullptr;
  const float4 *t4 = reinterpret_cast<const float4 *>(tri);
  uint32_t *keys = S.keys_a;
  uint32_t *tmp = S.keys_b;
  uint32_t *skeys = sorted_keys_out ? sorted_keys_out : keys;
  // an odd number of passes ends in the other buffer: sort into keys_b
  if (sort && sort_passes(bits) % 2 && skeys == keys) std::swap(skeys, tmp);
  int32_t *perm = perm_out ? perm_out : S.idx_a;
  MortonParams p{};
  uint32_t parts = 0;  // bounds partials the Morton pass folds itself
  if (bounds_host) {
    p = morton_params(ctx, bounds_host);
  } else {
    p.div_mode = ctx->div_mode;
    StageScope sc(ctx, LBVH_STAGE_BOUNDS);
    parts = launch_bounds(t4, n, S.partials, ctx->ctl, ctx->stream, S.cen,
                          use_cen ? kFoldPartialsMax : 0);
  }
  {
    StageScope sc(ctx, LBVH_STAGE_MORTON);
    uint32_t *cnt = sort ? S.digits : nullptr;
    if (use_cen)
      launch_morton_cen(S.cen, n, ctx->div_mode, keys, ctx->ctl, cnt, bits,
                        ctx->stream, S.partials, parts);
    else
      launch_aabb_morton(t4, n, p, bounds_host == nullptr, nullptr, keys,
                         ctx->ctl, ctx->stream, cnt, bits);
  }
  if (sort)
    run_sort(ctx, S.digits, keys, tmp, skeys, nullptr, S.idx_b, perm, n, true,
             bits);
  else
    launch_single_key(keys, skeys, perm, ctx->stream);
  run_tree(ctx, S.tree_items, S.tree_counts, tri, skeys, perm, n, tri_out,
           bvh_out, tile);
}

// Lays out one build's scratch from `base` (null: sizes only); returns the
// bytes it spans.  Regions start 256-byte aligned.
uint64_t layout_build_scratch(uint32_t n, int bits, int tile, bool use_cen,
                              char *base, BuildScratch *S) {
  uint64_t at = 0;
  const auto take = [&](uint64_t bytes) -> void * {
    void *p = base ? base + at : nullptr;
    at += (bytes + 255) & ~255ull;
    return p;
  };
  S->keys_a = static_cast<uint32_t *>(take(4ull * n));
  S->keys_b = static_cast<uint32_t *>(take(4ull * n));
  S->idx_a = static_cast<int32_t *>(take(4ull * n));
  S->idx_b = static_cast<int32_t *>(take(4ull * n));
  S->digits = static_cast<uint32_t *>(
      take(8ull * sort_tiles(n) * sort_radix(bits)));
  const uint64_t tiles = tree_tiles_total(n, tile);
  S->tree_items =
      static_cast<int4 *>(take(16ull * kItemInt4 * kTreeFrontier * tiles));
  S->tree_counts = static_cast<uint32_t *>(take(4ull * tiles));
  S->partials = static_cast<float4 *>(take(32ull * kBoundsBlocks));
  S->cen = use_cen ? static_cast<float *>(take(12ull * cen_stride(n))) : nullptr;
  return at;
}

// ---- batched builds ---------------------------------------------------------

// Builds (or reuses) the tile tables of a batch layout: sort tiles per mesh,
// and for every agglomeration level the tiles, item regions and counts of
// each mesh (the single-mesh level plan, per mesh).  Uploaded once.
int make_plan(lbvh_ctx *ctx, const uint32_t *offsets, uint32_t S) {
  auto &P = ctx->plan;
  uint32_t big = 0;
  for (uint32_t k = 0; k < S; ++k) big = std::max(big, offsets[k + 1] - offsets[k]);
  if (P.S == S && P.tile == tile_for(ctx, big) &&
      P.bits == bits_for(ctx, offsets[S]) &&
      std::equal(P.offsets.begin(), P.offsets.end(), offsets) &&
      P.offsets.size() == (size_t)S + 1)
    return LBVH_OK;
  drop_plan(ctx);
  P.offsets.assign(offsets, offsets + S + 1);
  P.S = S;
  uint32_t largest = 0;
  for (uint32_t k = 0; k < S; ++k)
    largest = std::max(largest, offsets[k + 1] - offsets[k]);
  P.tile = tile_for(ctx, largest);
  // one segmented sort runs over all the keys: its digit width is that of
  // a sort of the total
  P.bits = bits_for(ctx, offsets[S]);
  P.total = offsets[S];
  std::vector<uint32_t> tab;
  auto put = [&](const std::vector<uint32_t> &v) {
    const uint64_t at = tab.size();
    tab.insert(tab.end(), v.begin(), v.end());
    return at;
  };
  P.o_seg_off = put(P.offsets);
  std::vector<uint32_t> tile_seg, seg_tile0(S, 0);
  std::vector<int> levels_k(S, 0);
  std::vector<std::vector<uint32_t>> grid_k(S);
  for (uint32_t k = 0; k < S; ++k) {
    const uint32_t n = offsets[k + 1] - offsets[k];
    seg_tile0[k] = (uint32_t)tile_seg.size();
    for (uint32_t j = 0; j < batch_sort_tiles(n); ++j) tile_seg.push_back(k);
    if (n) {
      grid_k[k].resize(kMaxTreeLevels);
      levels_k[k] = tree_levels(n, P.tile, grid_k[k].data());
      P.levels = std::max(P.levels, levels_k[k]);
    }
  }
  P.sort_tiles = (uint32_t)tile_seg.size();
  P.o_tile_seg = put(tile_seg);
  P.o_seg_tile0 = put(seg_tile0);
  // Tree levels: each mesh's tiles at level L are consecutive; a tile of
  // level L >= 1 reads the frontiers of its mesh's level L-1 tiles.
  std::vector<uint32_t> prev_tile0(S, 0), prev_tiles(S, 0);
  for (int L = 0; L < P.levels; ++L) {
    std::vector<uint32_t> ltile_seg, lseg_tile0(S, 0), lprev_n(S, 0);
    for (uint32_t k = 0; k < S; ++k) {
      if (L >= levels_k[k]) continue;
      lseg_tile0[k] = (uint32_t)ltile_seg.size();
      for (uint32_t j = 0; j < grid_k[k][L]; ++j) ltile_seg.push_back(k);
      lprev_n[k] = L ? grid_k[k][L - 1] : 0u;
    }
    P.level_tiles.push_back((uint32_t)ltile_seg.size());
    P.o_ltile_seg.push_back(put(ltile_seg));
    P.o_lseg_tile0.push_back(put(lseg_tile0));
    P.o_lprev0.push_back(put(prev_tile0));
    P.o_lprev_n.push_back(put(lprev_n));
    prev_tile0 = lseg_tile0;
    P.tree_tiles += ltile_seg.size();
  }
  P.tree_tiles = std::max<uint64_t>(P.tree_tiles, 1);
  int st;
  if ((st = ensure(ctx, ctx->b_table, 4ull * tab.size(), false))) return st;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->b_table.ptr, tab.data(), 4ull * tab.size(),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LBVH_OK;
}

// Enqueues the whole batched pipeline (captured into a graph by the caller).
void enqueue_batch(lbvh_ctx *ctx, const float *tri, bool have_bounds,
                   void *bvh_out, float *tri_out, uint32_t *skeys,
                   int32_t *perm) {
  const auto &P = ctx->plan;
  const uint32_t *T = static_cast<const uint32_t *>(ctx->b_table.ptr);
  uint32_t *ord = static_cast<uint32_t *>(ctx->b_ord.ptr);
  const float4 *t4 = reinterpret_cast<const float4 *>(tri);
  const SortSegs ss{T + P.o_tile_seg, T + P.o_seg_tile0, T + P.o_seg_off};
  launch_batch_prologue(ord, P.S, !have_bounds, ctx->ctl, ctx->stream);
  float *cen = have_bounds ? nullptr : static_cast<float *>(ctx->b_cen.ptr);
  if (!have_bounds)
    launch_batch_bounds(t4, ss, P.sort_tiles, ord, cen, P.total, ctx->stream);
  uint32_t *keys = static_cast<uint32_t *>(ctx->b_keys_a.ptr);
  uint32_t *tmp = static_cast<uint32_t *>(ctx->b_keys_b.ptr);
  uint32_t *out = skeys ? skeys : keys;
  if (sort_passes(P.bits) % 2 && out == keys) std::swap(out, tmp);
  int32_t *vtmp = static_cast<int32_t *>(ctx->b_idx_b.ptr);
  int32_t *vout = perm ? perm : static_cast<int32_t *>(ctx->b_idx_a.ptr);
  launch_batch_morton(t4, cen, P.total, ss, P.sort_tiles,
                      have_bounds ? static_cast<const float *>(ctx->b_bounds.ptr)
                                  : nullptr,
                      ord, ctx->div_mode, keys,
                      static_cast<uint32_t *>(ctx->b_digits.ptr), P.bits,
                      ctx->stream);
  uint32_t *cnt = static_cast<uint32_t *>(ctx->b_digits.ptr);
  uint32_t *off = cnt + (size_t)sort_radix(P.bits) * P.sort_tiles;
  const SortPlan SP =
      sort_plan(P.bits, kKeyBits, keys, tmp, out, nullptr, vtmp, vout);
  for (int p = 0; p < SP.passes; ++p)
    launch_sort_pass(P.bits, p, SP.src[p], SP.vsrc[p], SP.dst[p], SP.vdst[p],
                     P.total, P.sort_tiles, cnt, off, p == 0, ctx->ctl,
                     ctx->stream, &ss, P.S);
  int4 *items = static_cast<int4 *>(ctx->b_tree_items.ptr);
  uint32_t *counts = static_cast<uint32_t *>(ctx->b_tree_counts.ptr);
  uint64_t base = 0, prev = 0;
  for (int L = 0; L < P.levels; ++L) {
    if (P.level_tiles[L] == 0) break;
    const TreeSegs ts{T + P.o_ltile_seg[L], T + P.o_lseg_tile0[L],
                      T + P.o_seg_off, T + P.o_lprev0[L], T + P.o_lprev_n[L]};
    launch_tree_level(L, P.tile, P.level_tiles[L], t4, out, vout, P.total,
                      reinterpret_cast<float4 *>(tri_out),
                      items + kItemInt4 * kTreeFrontier * prev, counts + prev, 0u,
                      items + kItemInt4 * kTreeFrontier * base, counts + base,
                      static_cast<float4 *>(bvh_out), ctx->ctl, ts,
                      ctx->stream);
    prev = base;
    base += P.level_tiles[L];
  }
}

}  // namespace

// ---------------------------------------------------------------------------

lbvh::Ctl *lbvh_context_ctl(lbvh_ctx *ctx) { return ctx ? ctx->ctl : nullptr; }

extern "C" {

const char *lbvh_status_string(int status) {
  switch (status) {
    case LBVH_OK: return "ok";
    case LBVH_ERR_INVALID_ARGUMENT: return "invalid argument";
    case LBVH_ERR_SIZE_MISMATCH: return "size mismatch";
    case LBVH_ERR_OUT_OF_BUDGET: return "out of HBM budget";
    case LBVH_ERR_OUT_OF_MEMORY: return "out of device memory";
    case LBVH_ERR_DEVICE: return "device error";
    case LBVH_ERR_NO_DEVICE: return "no such device";
    case LBVH_ERR_BUSY: return "busy";
    case LBVH_ERR_NOT_FOUND: return "not found";
    case LBVH_ERR_CANCELLED: return "cancelled";
    case LBVH_ERR_TIMEOUT: return "timed out";
    default: return "unknown status";
  }
}

int lbvh_create(int device_ordinal, uint64_t hbm_budget_bytes,
                lbvh_ctx **out_ctx) {
  if (!out_ctx) return LBVH_ERR_INVALID_ARGUMENT;
  *out_ctx = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device_ordinal < 0 ||
      device_ordinal >= count) {
    (void)hipGetLastError();
    return LBVH_ERR_NO_DEVICE;
  }
  lbvh_ctx *ctx = new (std::nothrow) lbvh_ctx();
  if (!ctx) return LBVH_ERR_OUT_OF_MEMORY;
  ctx->device = device_ordinal;
  ctx->budget = hbm_budget_bytes;
  if (hipSetDevice(device_ordinal) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
          hipSuccess) {
    delete ctx;
    return LBVH_ERR_DEVICE;
  }
  void *p = nullptr;
  int st = dev_alloc(ctx, sizeof(Ctl), &p);
  if (st != LBVH_OK ||
      hipMemsetAsync(p, 0, sizeof(Ctl), ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    if (p) (void)hipFree(p);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return st != LBVH_OK ? st : LBVH_ERR_DEVICE;
  }
  ctx->ctl = static_cast<Ctl *>(p);
  ctx->static_scratch = static_scratch_address(device_ordinal);
  *out_ctx = ctx;
  return LBVH_OK;
}

int lbvh_destroy(lbvh_ctx *ctx) {
  if (!ctx) return LBVH_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drop_plan(ctx);
    for (DevBuf *b : all_scratch(ctx)) dev_release(ctx, *b);
    for (auto &m : ctx->marks) {
      ctx->ev_pool.push_back(m.a);
      ctx->ev_pool.push_back(m.b);
    }
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    release_staging(ctx);
    ctx->user_allocs.for_each([](void *p, uint64_t) { (void)hipFree(p); });
    ctx->user_allocs.clear();
    (void)hipFree(ctx->ctl);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return LBVH_OK;
}

const char *lbvh_last_error(const lbvh_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

int lbvh_set_div_mode(lbvh_ctx *ctx, int div_mode) {
  CTX_GUARD(ctx);
  if (div_mode != LBVH_DIV_MUL_RCP && div_mode != LBVH_DIV_IEEE)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "unknown division mode");
  ctx->div_mode = div_mode;
  return LBVH_OK;
}

int lbvh_get_div_mode(const lbvh_ctx *ctx) {
  return ctx ? ctx->div_mode : -1;
}

int lbvh_set_tuning(lbvh_ctx *ctx, int param, int value) {
  CTX_GUARD(ctx);
  switch (param) {
    case LBVH_TUNE_TREE_TILE:
      if (value == 0) {
        ctx->tree_tile = kTreeTileDefault;
        ctx->tree_tile_fixed = false;
        return LBVH_OK;
      }
      if (value != 256 && value != 512 && value != 1024)
        return fail(ctx, LBVH_ERR_INVALID_ARGUMENT,
                    "tree tile must be 256, 512 or 1024");
      ctx->tree_tile = value;
      ctx->tree_tile_fixed = true;
      return LBVH_OK;
    case LBVH_TUNE_SORT_BITS:
      if (value == 0) {
        ctx->sort_bits = kRadixBits;
        ctx->sort_bits_fixed = false;
        return LBVH_OK;
      }
      if (value != kRadixBits && value != kWideBits)
        return fail(ctx, LBVH_ERR_INVALID_ARGUMENT,
                    "sort digit width must be 8 or 10");
      ctx->sort_bits = value;
      ctx->sort_bits_fixed = true;
      return LBVH_OK;
    default:
      return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "unknown tuning parameter");
  }
}

int lbvh_get_tuning(const lbvh_ctx *ctx, int param) {
  return lbvh_get_tuning_for(ctx, param, 0);
}

int lbvh_get_tuning_for(const lbvh_ctx *ctx, int param, uint32_t n) {
  if (!ctx) return -1;
  switch (param) {
    case LBVH_TUNE_TREE_TILE: return tile_for(ctx, n);
    case LBVH_TUNE_SORT_BITS: return bits_for(ctx, n);
    default: return -1;
  }
}

// A tuned variant must beat the untuned one by this fraction.
constexpr float kTuneMargin = 0.02f;

int lbvh_autotune(lbvh_ctx *ctx, uint32_t n_min, uint32_t n_max, int reps,
                  float *ms_out) {
  CTX_GUARD(ctx);
  int st;
  if (n_min == 0 || n_max < n_min || n_max > LBVH_MAX_TRIS || reps < 1)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "bad autotune range");
  static const int tiles[3] = {256, 512, kTreeTileDefault};
  static const int widths[2] = {kRadixBits, kWideBits};
  static_assert(kTreeTileDefault == 1024, "the untuned tile is variant 2");
  const int b0 = bucket_of(n_min), b1 = bucket_of(n_max);
  const uint32_t nbig = 1u << b1;
  void *tri = nullptr, *bvh = nullptr, *otri = nullptr;
  if ((st = dev_alloc(ctx, 48ull * nbig, &tri))) return st;
  if ((st = dev_alloc(ctx, 64ull * nbig, &bvh))) {
    dev_free_raw(ctx, tri, 48ull * nbig);
    return st;
  }
  if ((st = dev_alloc(ctx, 48ull * nbig, &otri))) {
    dev_free_raw(ctx, tri, 48ull * nbig);
    dev_free_raw(ctx, bvh, 64ull * nbig);
    return st;
  }
  hipEvent_t a = nullptr, b = nullptr;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const bool tile_was_fixed = ctx->tree_tile_fixed;
  const bool bits_were_fixed = ctx->sort_bits_fixed;
  ctx->tree_tile_fixed = true;  // force each variant while timing
  ctx->sort_bits_fixed = true;
  const int saved_tile = ctx->tree_tile, saved_bits = ctx->sort_bits;
  int out_i = 0;
  st = LBVH_OK;
  for (int k = b0; k <= b1 && st == LBVH_OK; ++k) {
    const uint32_t n = 1u << k;
    launch_synth_uniform(0x5EED0000ull + k, n, 0.01, static_cast<float *>(tri),
                         ctx->stream);
    float med_v[6];
    for (int v = 0; v < 6 && st == LBVH_OK; ++v) {
      ctx->tree_tile = tiles[v / 2];
      ctx->sort_bits = widths[v % 2];
      std::vector<float> ms;
      for (int r = -1; r < reps && st == LBVH_OK; ++r) {
        (void)hipEventRecord(a, ctx->stream);
        st = lbvh_build(ctx, static_cast<float *>(tri), n, nullptr, bvh,
                        static_cast<float *>(otri), nullptr, nullptr);
        (void)hipEventRecord(b, ctx->stream);
        if (st == LBVH_OK && hipEventSynchronize(b) == hipSuccess && r >= 0) {
          float t = 0.0f;
          (void)hipEventElapsedTime(&t, a, b);
          ms.push_back(t);
        }
      }
      float med = NAN;
      if (!ms.empty()) {
        std::sort(ms.begin(), ms.end());
        med = ms[ms.size() / 2];
      }
      med_v[v] = med;
      if (ms_out) ms_out[out_i] = med;
      ++out_i;
    }
    if (st == LBVH_OK) {
      // The untuned choice stays unless a variant beats it by more than
      // kTuneMargin: on these uniform meshes neighbouring variants are often
      // within run-to-run noise, and a choice made on noise does not carry to
      // other meshes (leaf tile 1024 built clustered-1m 4 % faster than 512,
      // which the uniform 2^20 mesh had picked by 0.3 %).
      const int v0 = 2 * 2 + (untuned_bits(n) == kWideBits ? 1 : 0);
      int pick = v0;
      for (int v = 0; v < 6; ++v)
        if (med_v[v] < med_v[pick]) pick = v;
      if (!(med_v[pick] < med_v[v0] * (1.0f - kTuneMargin))) pick = v0;
      ctx->tune_tile[k] = tiles[pick / 2];
      ctx->tune_bits[k] = widths[pick % 2];
    }
  }
  ctx->tree_tile = saved_tile;
  ctx->sort_bits = saved_bits;
  ctx->tree_tile_fixed = (st == LBVH_OK) ? false : tile_was_fixed;
  ctx->sort_bits_fixed = (st == LBVH_OK) ? false : bits_were_fixed;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  dev_free_raw(ctx, tri, 48ull * nbig);
  dev_free_raw(ctx, bvh, 64ull * nbig);
  dev_free_raw(ctx, otri, 48ull * nbig);
  return st;
}

void *lbvh_stream(lbvh_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int lbvh_synchronize(lbvh_ctx *ctx) {
  CTX_GUARD(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  uint32_t fault = 0;
  HIP_TRY(ctx, hipMemcpy(&fault, &ctx->ctl->fault, 4, hipMemcpyDeviceToHost));
  if (fault) {
    HIP_TRY(ctx, hipMemset(&ctx->ctl->fault, 0, 4));
    std::string what;
    if (fault & kFaultNotATree)
      what += "; the up pass's parents or index array is not a radix tree "
              "over the leaves (an entry out of range, a path longer than "
              "any tree's, a node with a third child, or no root)";
    if (fault & kFaultUpFrontier)
      what += "; an up pass block left more subtrees than a tree allows";
    if (fault & kFaultFrontier)
      what += "; a build tile's frontier overflowed";
    return fail(ctx, LBVH_ERR_DEVICE,
                "a tree kernel found inconsistent input" + what +
                    "; results of the operations since the last synchronize "
                    "are invalid");
  }
  return LBVH_OK;
}

uint64_t lbvh_bytes_in_use(const lbvh_ctx *ctx) {
  return ctx ? ctx->in_use : 0;
}

int lbvh_trim(lbvh_ctx *ctx) {
  CTX_GUARD(ctx);
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  drop_plan(ctx);
  for (DevBuf *b : all_scratch(ctx)) dev_release(ctx, *b);
  return LBVH_OK;
}

int lbvh_alloc(lbvh_ctx *ctx, uint64_t bytes, void **out_dptr) {
  CTX_GUARD(ctx);
  if (!out_dptr) return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null output");
  *out_dptr = nullptr;
  void *p = nullptr;
  int st = dev_alloc(ctx, bytes, &p);
  if (st != LBVH_OK) return st;
  ctx->user_allocs.add(p, bytes);
  *out_dptr = p;
  return LBVH_OK;
}

int lbvh_free(lbvh_ctx *ctx, void *dptr) {
  CTX_GUARD(ctx);
  if (!dptr) return LBVH_OK;
  if (ctx->user_allocs.check(dptr, 0) == AllocTable::kForeign)
    return fail(ctx, LBVH_ERR_NOT_FOUND, "pointer was not allocated by lbvh_alloc");
  uint64_t bytes = 0;
  if (!ctx->user_allocs.find(dptr, &bytes))
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT,
                "pointer lies inside an lbvh_alloc buffer but is not its start");
  // Work already queued may still read the buffer.
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipFree(dptr));
  ctx->in_use -= bytes;
  ctx->user_allocs.remove(dptr, nullptr);
  return LBVH_OK;
}

int lbvh_upload(lbvh_ctx *ctx, void *dst_dev, const void *src_host,
                uint64_t bytes) {
  CTX_GUARD(ctx);
  if (bytes == 0) return LBVH_OK;
  if (!dst_dev || !src_host)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  int st;
  if ((st = check_spans(ctx, {{dst_dev, bytes, "upload destination"}})))
    return st;
  if (bytes >= kStageMin && ensure_staging(ctx)) {
    HIP_TRY(ctx, staged_copy(ctx, static_cast<char *>(dst_dev),
                             static_cast<const char *>(src_host), bytes, true));
    return LBVH_OK;
  }
  HIP_TRY(ctx, hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LBVH_OK;
}

int lbvh_download(lbvh_ctx *ctx, void *dst_host, const void *src_dev,
                  uint64_t bytes) {
  CTX_GUARD(ctx);
  if (bytes == 0) return LBVH_OK;
  if (!dst_host || !src_dev)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  int st;
  if ((st = check_spans(ctx, {{src_dev, bytes, "download source"}})))
    return st;
  if (bytes >= kStageMin && ensure_staging(ctx)) {
    HIP_TRY(ctx, staged_copy(ctx, static_cast<char *>(dst_host),
                             static_cast<const char *>(src_dev), bytes, false));
    return LBVH_OK;
  }
  HIP_TRY(ctx, hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LBVH_OK;
}

// ---- per-stage entry points ----------------------------------------------

int lbvh_bounds(lbvh_ctx *ctx, const float *tri, uint32_t n,
                float *bounds_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (!bounds_out || (n && !tri))
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if (n == 0)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "bounds of an empty mesh");
  if ((st = check_spans(ctx, {{tri, 48ull * n, "triangles"},
                              {bounds_out, 24, "bounds output"}})))
    return st;
  if ((st = ensure_partials(ctx))) return st;
  launch_bounds(reinterpret_cast<const float4 *>(tri), n,
                static_cast<float4 *>(ctx->partials.ptr), ctx->ctl,
                ctx->stream);
  launch_copy_bounds(ctx->ctl, bounds_out, ctx->stream);
  return check_launch(ctx, "lbvh_bounds");
}

int lbvh_aabb_morton(lbvh_ctx *ctx, const float *tri, uint32_t n,
                     const float *bounds_host, float *aabb_out,
                     uint32_t *keys_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (n == 0) return LBVH_OK;
  if (!tri || !keys_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{tri, 48ull * n, "triangles"},
                              {aabb_out, 32ull * n, "AABB output"},
                              {keys_out, 4ull * n, "key output"}})))
    return st;
  if ((st = check_bounds(ctx, bounds_host, 1))) return st;
  MortonParams p{};
  if (bounds_host) {
    p = morton_params(ctx, bounds_host);
  } else {
    if ((st = ensure_partials(ctx))) return st;
    p.div_mode = ctx->div_mode;
    launch_bounds(reinterpret_cast<const float4 *>(tri), n,
                  static_cast<float4 *>(ctx->partials.ptr), ctx->ctl,
                  ctx->stream);
  }
  launch_aabb_morton(reinterpret_cast<const float4 *>(tri), n, p,
                     bounds_host == nullptr,
                     reinterpret_cast<float4 *>(aabb_out), keys_out, ctx->ctl,
                     ctx->stream);
  return check_launch(ctx, "lbvh_aabb_morton");
}

int lbvh_sort(lbvh_ctx *ctx, uint32_t *keys, int32_t *idx_out, uint32_t n) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (n == 0) return LBVH_OK;
  if (!keys || !idx_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{keys, 4ull * n, "keys"},
                              {idx_out, 4ull * n, "index output"}})))
    return st;
  if (n == 1) {
    launch_single_key(keys, keys, idx_out, ctx->stream);
    return check_launch(ctx, "lbvh_sort");
  }
  // Any u32 key, as the reference's 16 passes of 2 bits sort: all 32 bits.
  // 10-bit digits take 4 passes here too, so they run only when fixed.
  const int bits = ctx->sort_bits_fixed ? ctx->sort_bits : kRadixBits;
  if ((st = ensure_sort_scratch(ctx, n, bits))) return st;
  run_sort(ctx, static_cast<uint32_t *>(ctx->digits.ptr), keys,
           static_cast<uint32_t *>(ctx->keys_b.ptr), keys,
           static_cast<uint32_t *>(ctx->keys_a.ptr),
           static_cast<int32_t *>(ctx->idx_a.ptr), idx_out, n, false, bits,
           kFullKeyBits);
  return check_launch(ctx, "lbvh_sort");
}

int lbvh_radix_tree(lbvh_ctx *ctx, const uint32_t *sorted_keys, uint32_t n,
                    int32_t *parents_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (n == 0) return LBVH_OK;
  if (!sorted_keys || !parents_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{sorted_keys, 4ull * n, "sorted keys"},
                              {parents_out, 8ull * n, "parents output"}})))
    return st;
  if ((st = ensure(ctx, ctx->rt_scratch, radix_tree_scratch_bytes(n), false)))
    return st;
  launch_radix_tree(sorted_keys, n, parents_out, ctx->rt_scratch.ptr,
                    ctx->stream);
  return check_launch(ctx, "lbvh_radix_tree");
}

int lbvh_up_pass(lbvh_ctx *ctx, const int32_t *idx, const float *aabb,
                 const int32_t *parents, uint32_t n, void *bvh_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (n == 0) return LBVH_OK;
  if (!idx || !aabb || !parents || !bvh_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{idx, 4ull * n, "index"},
                              {aabb, 32ull * n, "AABBs"},
                              {parents, 8ull * n, "parents"},
                              {bvh_out, 64ull * n, "BVH output"}})))
    return st;
  if ((st = ensure(ctx, ctx->up_items, up_pass_scratch_bytes(n), false)))
    return st;
  launch_up_pass(idx, reinterpret_cast<const float4 *>(aabb), parents, n,
                 static_cast<float4 *>(bvh_out), ctx->up_items.ptr, ctx->ctl,
                 ctx->stream);
  return check_launch(ctx, "lbvh_up_pass");
}

int lbvh_rearrange(lbvh_ctx *ctx, const float *tri_in, const int32_t *idx,
                   uint32_t n, float *tri_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (n == 0) return LBVH_OK;
  if (!tri_in || !idx || !tri_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{tri_in, 48ull * n, "triangles"},
                              {idx, 4ull * n, "index"},
                              {tri_out, 48ull * n, "triangle output"}})))
    return st;
  launch_rearrange(reinterpret_cast<const float4 *>(tri_in), idx, n,
                   reinterpret_cast<float4 *>(tri_out), ctx->stream);
  return check_launch(ctx, "lbvh_rearrange");
}

// ---- fused build / refit ----------------------------------------------------

int lbvh_build(lbvh_ctx *ctx, const float *tri, uint32_t n,
               const float *bounds_host, void *bvh_out, float *tri_out,
               uint32_t *sorted_keys_out, int32_t *perm_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_build_args(ctx, tri, n, bvh_out, tri_out, sorted_keys_out,
                             perm_out)) || n == 0)
    return st;
  if ((st = check_bounds(ctx, bounds_host, 1))) return st;
  const int bits = bits_for(ctx, n);
  const int tile = tile_for(ctx, n);
  const bool use_cen = !bounds_host && n > 1;
  if ((st = ensure_sort_scratch(ctx, n, bits))) return st;
  if ((st = ensure_tree_scratch(ctx, n, tile))) return st;
  if (!bounds_host && (st = ensure_partials(ctx))) return st;
  if (use_cen && (st = ensure(ctx, ctx->cen, 12ull * cen_stride(n), false)))
    return st;
  BuildScratch S;
  S.keys_a = static_cast<uint32_t *>(ctx->keys_a.ptr);
  S.keys_b = static_cast<uint32_t *>(ctx->keys_b.ptr);
  S.idx_a = static_cast<int32_t *>(ctx->idx_a.ptr);
  S.idx_b = static_cast<int32_t *>(ctx->idx_b.ptr);
  S.digits = static_cast<uint32_t *>(ctx->digits.ptr);
  S.tree_items = static_cast<int4 *>(ctx->tree_items.ptr);
  S.tree_counts = static_cast<uint32_t *>(ctx->tree_counts.ptr);
  S.partials = static_cast<float4 *>(ctx->partials.ptr);
  S.cen = use_cen ? static_cast<float *>(ctx->cen.ptr) : nullptr;
  build_pipeline(ctx, S, tri, n, bounds_host, bvh_out, tri_out,
                 sorted_keys_out, perm_out, bits, tile);
  return check_launch(ctx, "lbvh_build");
}

int lbvh_build_static(lbvh_ctx *ctx, const float *tri, uint32_t n,
                      const float *bounds_host, void *bvh_out, float *tri_out,
                      uint32_t *sorted_keys_out, int32_t *perm_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_build_args(ctx, tri, n, bvh_out, tri_out, sorted_keys_out,
                             perm_out)) || n == 0)
    return st;
  if ((st = check_bounds(ctx, bounds_host, 1))) return st;
  if (n > kStaticMaxTris)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT,
                "lbvh_build_static takes at most " +
                    std::to_string(kStaticMaxTris) + " triangles (" +
                    std::to_string(n) + " given); use lbvh_build");
  if (!ctx->static_scratch)
    return fail(ctx, LBVH_ERR_DEVICE,
                "the static scratch has no address on this device");
  const int bits = bits_for(ctx, n);
  const int tile = tile_for(ctx, n);
  BuildScratch S;
  const uint64_t need =
      layout_build_scratch(n, bits, tile, !bounds_host && n > 1,
                           static_cast<char *>(ctx->static_scratch), &S);
  if (need > kStaticScratchBytes)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT,
                "the build needs " + std::to_string(need) +
                    " bytes of scratch, more than the static " +
                    std::to_string(kStaticScratchBytes));
  // no profiling events: they would be created here on first use
  const bool prof = ctx->prof;
  ctx->prof = false;
  build_pipeline(ctx, S, tri, n, bounds_host, bvh_out, tri_out,
                 sorted_keys_out, perm_out, bits, tile);
  ctx->prof = prof;
  return check_launch(ctx, "lbvh_build_static");
}

int lbvh_refit(lbvh_ctx *ctx, const float *tri, uint32_t n,
               const uint32_t *sorted_keys, const int32_t *perm, void *bvh_out,
               float *tri_out) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (n == 0) return LBVH_OK;
  if (!tri || !sorted_keys || !perm || !bvh_out || !tri_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{tri, 48ull * n, "triangles"},
                              {sorted_keys, 4ull * n, "sorted keys"},
                              {perm, 4ull * n, "permutation"},
                              {bvh_out, 64ull * n, "BVH output"},
                              {tri_out, 48ull * n, "triangle output"}})))
    return st;
  const int tile = tile_for(ctx, n);
  if ((st = ensure_tree_scratch(ctx, n, tile))) return st;
  run_tree(ctx, static_cast<int4 *>(ctx->tree_items.ptr),
           static_cast<uint32_t *>(ctx->tree_counts.ptr), tri, sorted_keys,
           perm, n, tri_out, bvh_out, tile);
  return check_launch(ctx, "lbvh_refit");
}


// ---- batched build -------------------------------------------------------------

int lbvh_build_batched(lbvh_ctx *ctx, const float *tri,
                       const uint32_t *offsets, uint32_t num_meshes,
                       const float *bounds_host, void *bvh_out, float *tri_out,
                       uint32_t *sorted_keys_out, int32_t *perm_out) {
  CTX_GUARD(ctx);
  int st;
  if (!offsets) return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null offsets");
  if (num_meshes == 0) return LBVH_OK;
  if (offsets[0] != 0)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "offsets[0] must be 0");
  for (uint32_t k = 0; k < num_meshes; ++k)
    if (offsets[k + 1] < offsets[k])
      return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "offsets must not decrease");
  const uint32_t total = offsets[num_meshes];
  if ((st = check_n(ctx, total))) return st;
  if (total == 0) return LBVH_OK;
  if (!tri || !bvh_out || !tri_out)
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{tri, 48ull * total, "triangles"},
                              {bvh_out, 64ull * total, "BVH output"},
                              {tri_out, 48ull * total, "triangle output"},
                              {sorted_keys_out, 4ull * total, "sorted-key output"},
                              {perm_out, 4ull * total, "permutation output"}})))
    return st;
  if ((st = check_bounds(ctx, bounds_host, num_meshes))) return st;
  if ((st = make_plan(ctx, offsets, num_meshes))) return st;
  const auto &P = ctx->plan;
  if ((st = ensure(ctx, ctx->b_keys_a, 4ull * total, false))) return st;
  if ((st = ensure(ctx, ctx->b_keys_b, 4ull * total, false))) return st;
  if ((st = ensure(ctx, ctx->b_idx_a, 4ull * total, false))) return st;
  if ((st = ensure(ctx, ctx->b_idx_b, 4ull * total, false))) return st;
  const uint64_t radix = sort_radix(P.bits);
  if ((st = ensure(ctx, ctx->b_digits, 8ull * radix * P.sort_tiles, false)))
    return st;
  if ((st = ensure(ctx, ctx->b_ord, 32ull * num_meshes, false))) return st;
  if ((st = ensure(ctx, ctx->b_tree_items,
                   16ull * kItemInt4 * kTreeFrontier * P.tree_tiles, false)))
    return st;
  if ((st = ensure(ctx, ctx->b_tree_counts, 4ull * P.tree_tiles, false)))
    return st;
  if (!bounds_host &&
      (st = ensure(ctx, ctx->b_cen, 12ull * cen_stride(total), false)))
    return st;
  if (bounds_host) {
    if ((st = ensure(ctx, ctx->b_bounds, 24ull * num_meshes, false))) return st;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->b_bounds.ptr, bounds_host,
                                24ull * num_meshes, hipMemcpyHostToDevice,
                                ctx->stream));
  }
  // Replay the captured graph when the batch layout, buffers and options
  // match the last call; otherwise capture the pipeline anew.
  std::vector<const void *> key = {
      tri, bvh_out, tri_out, sorted_keys_out, perm_out,
      reinterpret_cast<const void *>((uintptr_t)(bounds_host ? 1 : 0)),
      reinterpret_cast<const void *>((uintptr_t)ctx->div_mode),
      reinterpret_cast<const void *>((uintptr_t)P.bits)};
  for (DevBuf *b : all_scratch(ctx)) key.push_back(b->ptr);
  auto &Pm = ctx->plan;
  if (!Pm.exec || Pm.graph_key != key) {
    if (Pm.exec) (void)hipGraphExecDestroy(Pm.exec);
    if (Pm.graph) (void)hipGraphDestroy(Pm.graph);
    Pm.exec = nullptr;
    Pm.graph = nullptr;
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream,
                                       hipStreamCaptureModeThreadLocal));
    enqueue_batch(ctx, tri, bounds_host != nullptr, bvh_out, tri_out,
                  sorted_keys_out, perm_out);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
    Pm.graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&Pm.exec, g, nullptr, nullptr, 0));
    Pm.graph_key = key;
  }
  HIP_TRY(ctx, hipGraphLaunch(Pm.exec, ctx->stream));
  return check_launch(ctx, "lbvh_build_batched");
}

// ---- consumer check ------------------------------------------------------------

int lbvh_raycast(lbvh_ctx *ctx, const void *bvh, const float *tri, uint32_t n,
                 const float *rays, uint32_t count, float *hits) {
  CTX_GUARD(ctx);
  int st;
  if ((st = check_n(ctx, n))) return st;
  if (count == 0) return LBVH_OK;
  if (!rays || !hits || (n && (!bvh || !tri)))
    return fail(ctx, LBVH_ERR_INVALID_ARGUMENT, "null pointer");
  if ((st = check_spans(ctx, {{bvh, 64ull * n, "BVH"},
                              {tri, 48ull * n, "triangles"},
                              {rays, 32ull * count, "rays"},
                              {hits, 16ull * count, "hit output"}})))
    return st;
  if (n == 0) {
    HIP_TRY(ctx, hipMemsetAsync(hits, 0, 16ull * count, ctx->stream));
    return LBVH_OK;
  }
  launch_raycast(bvh, tri, n, rays, count, hits, ctx->stream);
  return check_launch(ctx, "lbvh_raycast");
}

// ---- profiling ---------------------------------------------------------------

int lbvh_profile_enable(lbvh_ctx *ctx, int enable) {
  CTX_GUARD(ctx);
  ctx->prof = enable != 0;
  return LBVH_OK;
}

int lbvh_profile_read(lbvh_ctx *ctx, lbvh_profile *out, int reset) {
  CTX_GUARD(ctx);
  if (!ctx->marks.empty()) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &m : ctx->marks) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
        ctx->acc.ms[m.stage] += ms;
        ctx->acc.launches[m.stage] += 1;
      }
      ctx->ev_pool.push_back(m.a);
      ctx->ev_pool.push_back(m.b);
    }
    ctx->marks.clear();
  }
  if (out) *out = ctx->acc;
  if (reset) ctx->acc = lbvh_profile{};
  return LBVH_OK;
}

}  // extern "C"
