// server.hip -- device-resident LBVH build/refit service (lbvh_server_*).
//
// Continuous batching over execution slots: each slot is an lbvh_ctx with
// its own HIP stream (default priority) and scratch, so requests admitted to
// different slots overlap on the GPU.  A dispatcher thread admits queued
// requests in arrival order as slots and HBM budget free up (requests on a
// mesh that is busy, or that do not fit yet, wait without blocking refits
// behind them), watches completion events, and evicts least recently used
// idle meshes when a new build needs room.  Each slot has a worker thread
// that enqueues its requests' device work on the slot's stream, up to
// kSlotDepth requests ahead, so host-side launch costs overlap across slots
// and with the GPU.  A worker only ever takes its own slot's lock (never the
// server lock, which the dispatcher holds for most of each round), so it is
// never queued behind admission work.  Lock order: server `mu`, then a
// slot's `m`, then `done_mu`.  Nothing here waits on the GPU except
// lbvh_server_drain.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#define LBVH_HD __host__ __device__
#include "../../include/lbvh.h"
#include "lbvh_internal.h"
#include "range_arena.h"
#include "synth_math.h"

namespace lbvh {
void launch_synth_uniform(uint64_t seed, uint32_t n, double half_extent,
                          float *tri, hipStream_t s);
void launch_synth_jitter(uint64_t seed, uint32_t n, double sigma, float *tri,
                         hipStream_t s);
}  // namespace lbvh

namespace {

// ---- per-request epilogue -----------------------------------------------------

// What a request's device work leaves for the dispatcher, in pinned host
// memory: the slot context's fault bits and the result digests.
struct RequestResult {
  uint32_t fault;
  uint32_t pad;
  unsigned long long digest[2];  // BVH words, O_TRI words
};

__device__ __forceinline__ unsigned long long digest_weight(uint64_t j) {
  return lbvh_synth::splitmix64(j) | 1ull;
}

__device__ __forceinline__ unsigned long long digest4(uint4 v, uint64_t j) {
  return (unsigned long long)v.x * digest_weight(j) +
         (unsigned long long)v.y * digest_weight(j + 1) +
         (unsigned long long)v.z * digest_weight(j + 2) +
         (unsigned long long)v.w * digest_weight(j + 3);
}

constexpr int kDigestThreads = 256;

// sum[0] += the BVH words' weighted sum, sum[1] += the O_TRI words' (both
// modulo 2^64, so the order of the additions does not matter).
__global__ __launch_bounds__(kDigestThreads) void k_digest(
    const uint4 *__restrict__ bvh, const uint4 *__restrict__ otri, uint32_t n,
    unsigned long long *sum) {
  __shared__ unsigned long long s_part[2][kDigestThreads / 64];
  const uint64_t nb = 4ull * n, no = 3ull * n;  // 16-byte vectors
  unsigned long long a = 0, b = 0;
  const uint64_t stride = (uint64_t)gridDim.x * kDigestThreads;
  for (uint64_t q = (uint64_t)blockIdx.x * kDigestThreads + threadIdx.x;
       q < nb; q += stride) {
    a += digest4(bvh[q], 4 * q);
    if (q < no) b += digest4(otri[q], 4 * q);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_part[0][w] = a;
    s_part[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kDigestThreads / 64; ++k) {
      a += s_part[0][k];
      b += s_part[1][k];
    }
    atomicAdd(&sum[0], a);
    atomicAdd(&sum[1], b);
  }
}

// After a request's kernels: hands the fault bits and digests to the host
// and clears them for the slot's next request (same stream, so they hold
// exactly this request's).
__global__ void k_request_done(lbvh::Ctl *ctl, unsigned long long *sum,
                               RequestResult *out) {
  out->fault = ctl->fault;
  out->digest[0] = sum[0];
  out->digest[1] = sum[1];
  ctl->fault = 0u;
  sum[0] = 0ull;
  sum[1] = 0ull;
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Bytes a resident mesh holds: triangles 48, BVH 64, O_TRI 48, sorted keys
// 4, permutation 4 per triangle.
constexpr uint64_t kMeshBytesPerTri = 168;
// Scratch a slot may need to grow to for a build / refit.  A build's
// buffers: keys and indices 16, box centres 12, sort tile counts and
// offsets 2 (8 for sorts of at most 2^20 keys, whose tiles hold 1,024
// keys), agglomeration item slots up to 27.5 (128 x 48 B per 256-leaf tile,
// all levels): at most 63.5.  A refit's: the item slots, 27.5.  A context
// grows a buffer to 1.25 x what it is asked for (ensure), so 79.4 and 34.4.
constexpr uint64_t kBuildScratchPerTri = 80;
constexpr uint64_t kRefitScratchPerTri = 36;
// plus what does not scale with n: the context's control block, the bounds
// partials (80 KB with headroom), and tile tables rounded up to whole tiles
constexpr uint64_t kSlotScratchFixed = 1ull << 20;

struct Mesh {
  uint32_t n = 0;
  void *buf = nullptr;  // one arena range: tri | bvh | otri | keys | perm
  uint64_t bytes = 0;   // the range's size
  uint32_t chunk = 0;   // the arena chunk holding it
  bool valid = false;   // holds a completed build
  bool busy = false;    // a request on it is queued in a slot
  uint64_t last_use = 0;
  float *tri() const { return static_cast<float *>(buf); }
  void *bvh() const { return static_cast<char *>(buf) + 48ull * n; }
  float *otri() const {
    return reinterpret_cast<float *>(static_cast<char *>(buf) + 112ull * n);
  }
  uint32_t *keys() const {
    return reinterpret_cast<uint32_t *>(static_cast<char *>(buf) + 160ull * n);
  }
  int32_t *perm() const {
    return reinterpret_cast<int32_t *>(static_cast<char *>(buf) + 164ull * n);
  }
};

// Mesh memory.  Device chunks, allocated as the resident set grows, are
// carved by a best-fit range allocator that merges neighbouring free ranges.
// A mesh is freed only when idle (its last request completed), so its range
// is reusable at once by any slot, and admission in steady state makes no
// driver call: a driver pool that keeps freed blocks cannot always reuse one
// for another size, and once those fill HBM each allocation failure costs a
// trim (thousands of unmaps) and fresh mappings while admission waits.
struct HipChunks {
  static void *alloc(uint64_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    return p;
  }
  static void release(void *p) { (void)hipFree(p); }  // waits for the device
};
using MeshArena = lbvh::RangeArena<HipChunks>;

struct Request {
  uint64_t id = 0, mesh_id = 0, seed = 0;
  int kind = 0, source = 0;
  uint32_t n = 0;
  double param = 0.0, submit_s = 0.0;
  double deadline_s = 0.0;  // admit by this time or time out (0: no limit)
  const float *dev = nullptr;
  std::vector<float> host;
};

// A request handed to a slot.  The dispatcher fills in the mesh buffers;
// the slot's worker enqueues the device work and records `ev`.
struct Work {
  Request req;
  Mesh mesh;          // copy of the mesh's buffer layout for the worker
  hipEvent_t ev = nullptr;
  int ring = 0;       // the slot's event and result entry of this request
  bool digest = false;
  double start_s = 0.0;
  bool launched = false;  // the worker has enqueued it (or failed to)
  int status = LBVH_OK;   // the enqueue's status
};

constexpr int kSlotDepth = 2;  // requests in flight per slot
// How long an idle slot worker polls its mailbox before it blocks.
constexpr double kWorkerPollS = 200e-6;
// Admission looks at this many requests at the head of the queue per round:
// enough to find runnable work past held ones, and it keeps a round's cost
// independent of a backlog (a scan of the whole queue made each round slower
// as the queue grew, which grew the queue).
constexpr size_t kAdmitWindow = 64;

struct Slot {
  lbvh_ctx *ctx = nullptr;
  // lbvh_bytes_in_use(ctx) as of the worker's last enqueue (or a trim): the
  // dispatcher reads this instead of the context, whose lock the worker
  // holds while it enqueues
  std::atomic<uint64_t> scratch_bytes{0};
  // dispatcher only: a bound on the scratch the slot's context may hold,
  // from the largest build and the largest refit admitted to it since its
  // scratch was last released.  The context keeps every buffer at the
  // largest size any request asked of it, and a build and a refit use
  // different buffers, so the bound adds the two estimates; it also covers
  // growth admitted but not yet allocated, which must not be handed out
  // twice.
  uint64_t max_build_n = 0, max_refit_n = 0;
  uint64_t bound = 0;
  hipEvent_t events[kSlotDepth] = {};
  uint32_t next_event = 0;  // dispatcher only
  // per event: what the request's epilogue left (pinned host memory, and
  // its device address), and the device words its digest accumulates in
  RequestResult *results = nullptr;
  RequestResult *results_dev = nullptr;
  unsigned long long *sums = nullptr;
  std::mutex m;             // guards mailbox, inflight and their Work
  std::deque<Work> mailbox;   // assigned, not yet taken by the worker
  std::atomic<uint32_t> mail_n{0};  // mailbox.size(), readable without m
  std::deque<Work> inflight;  // taken by the worker, in stream order
  std::condition_variable wcv;
  std::thread worker;
  bool idle() {
    std::lock_guard<std::mutex> g(m);
    return mailbox.empty() && inflight.empty();
  }
  size_t depth() {
    std::lock_guard<std::mutex> g(m);
    return mailbox.size() + inflight.size();
  }
};

}  // namespace

struct lbvh_server {
  int device = 0;
  uint64_t budget = 0;
  int div_mode = LBVH_DIV_MUL_RCP;
  bool digest = false;  // admitted requests digest their results
  std::atomic<double> max_wait{0.0};  // queueing limit of new requests
  std::deque<Slot> slots;  // stable addresses: workers hold references
  MeshArena arena;
  std::mutex mu;
  std::condition_variable cv;       // wakes the dispatcher
  std::condition_variable idle_cv;  // wakes lbvh_server_drain
  std::deque<Request> queue;
  // Submissions and completions cross their own small locks, so submit and
  // poll never wait for the dispatcher's admission work under `mu`.
  std::mutex in_mu;
  std::deque<Request> inbox;
  std::atomic<uint32_t> inbox_n{0};
  std::mutex done_mu;
  std::deque<lbvh_completion> done;
  std::unordered_map<uint64_t, Mesh> meshes;
  uint64_t next_id = 1, tick = 0, resident_bytes = 0, peak = 0;
  lbvh_server_stats st{};
  std::string last_error;
  std::atomic<bool> stop{false};
  std::thread thr;
};

namespace {

uint64_t slots_bytes(lbvh_server *s) {
  uint64_t b = 0;
  for (auto &sl : s->slots) b += std::max<uint64_t>(sl.scratch_bytes, sl.bound);
  return b;
}

// Device memory held: every arena chunk (free ranges included) and every
// slot's scratch.  The budget bounds this.
uint64_t in_use(lbvh_server *s) {
  return s->arena.reserved() + slots_bytes(s);
}

bool all_idle(lbvh_server *s) {
  for (auto &sl : s->slots)
    if (!sl.idle()) return false;
  return true;
}

// A mesh is only freed while idle (its last request completed), so its
// range goes back to the arena at once.
void free_mesh(lbvh_server *s, Mesh &m) {
  if (m.buf) {
    s->arena.free({static_cast<char *>(m.buf), m.chunk, m.bytes});
    s->resident_bytes -= m.bytes;
  }
  m.buf = nullptr;
  m.bytes = 0;
  m.n = 0;
  m.valid = false;
}

// Memory a request still needs: a new mesh range (0 if its mesh already has
// one of its size) and slot scratch beyond what the slot holds.
struct Need {
  uint64_t mesh = 0, scratch = 0;
  uint64_t slot_total = 0;  // the slot's scratch bound with this request
  uint64_t max_build_n = 0, max_refit_n = 0;  // the slot's, with it
  bool any() const { return mesh || scratch; }
};
Need request_need(lbvh_server *s, const Request &r, const Slot &sl) {
  Need need;
  if (r.kind == LBVH_REQ_BUILD) {
    auto it = s->meshes.find(r.mesh_id);
    if (it == s->meshes.end() || !it->second.buf || it->second.n != r.n)
      need.mesh = MeshArena::round(kMeshBytesPerTri * r.n);
  }
  need.max_build_n = sl.max_build_n;
  need.max_refit_n = sl.max_refit_n;
  if (r.kind == LBVH_REQ_BUILD)
    need.max_build_n = std::max<uint64_t>(need.max_build_n, r.n);
  else
    need.max_refit_n = std::max<uint64_t>(need.max_refit_n, r.n);
  const uint64_t bound = kBuildScratchPerTri * need.max_build_n +
                         kRefitScratchPerTri * need.max_refit_n +
                         kSlotScratchFixed;
  const uint64_t have = std::max<uint64_t>(sl.scratch_bytes, sl.bound);
  if (bound > have) need.scratch = bound - have;
  need.slot_total = std::max(bound, have);
  return need;
}

bool evict_lru(lbvh_server *s, uint64_t keep) {
  uint64_t victim = 0, oldest = UINT64_MAX;
  for (auto &kv : s->meshes)
    if (kv.first != keep && !kv.second.busy && kv.second.buf &&
        kv.second.last_use < oldest) {
      oldest = kv.second.last_use;
      victim = kv.first;
    }
  if (oldest == UINT64_MAX) return false;
  free_mesh(s, s->meshes[victim]);
  s->meshes.erase(victim);
  s->st.evictions += 1;
  return true;
}

// Frees idle memory until request r fits on slot sl: a free arena range
// holds its mesh (or a new chunk for it fits the budget) and its slot
// scratch fits the budget.  Chunks left empty go first when the budget is
// short, then least recently used idle meshes other than r's, then idle
// slots' scratch.  The need is recomputed after every step (a trim of sl
// itself changes it); *out gets the need the request fits with.  True if it
// fits.
bool make_room(lbvh_server *s, const Request &r, Slot &target, Need *out) {
  const uint64_t keep = r.mesh_id;
  while (true) {
    const Need need = request_need(s, r, target);
    *out = need;
    if (!need.any()) return true;
    const bool fit = !need.mesh || s->arena.can_fit(need.mesh);
    const uint64_t grow =
        fit ? 0 : MeshArena::chunk_for(need.mesh, s->budget);
    const bool in_budget =
        !s->budget || in_use(s) + grow + need.scratch <= s->budget;
    if (in_budget) {
      if (fit) return true;
      if (s->arena.add_chunk(grow)) continue;  // the device is full if not
    } else if (s->arena.release_empty()) {
      continue;
    }
    if (evict_lru(s, keep)) continue;
    if (s->arena.release_empty()) continue;
    // A trim is progress only if it released bytes or a bound, so the loop
    // ends even if a context keeps more than its control block.
    bool trimmed = false;
    for (auto &sl : s->slots)
      if (sl.idle() && (sl.scratch_bytes > sizeof(lbvh::Ctl) || sl.bound)) {
        // an idle slot's worker waits for work that only this thread
        // hands out, so it does not touch the context meanwhile
        const uint64_t before = sl.scratch_bytes;
        const bool bounded = sl.bound != 0;
        lbvh_trim(sl.ctx);
        sl.scratch_bytes = lbvh_bytes_in_use(sl.ctx);
        sl.bound = 0;
        sl.max_build_n = sl.max_refit_n = 0;
        if (bounded || sl.scratch_bytes < before) trimmed = true;
      }
    if (!trimmed) return false;
  }
}

void complete(lbvh_server *s, const Request &r, int status, uint32_t slot,
              double start, const RequestResult *res = nullptr) {
  lbvh_completion c{};
  if (res) {
    c.digest_bvh = res->digest[0];
    c.digest_tri = res->digest[1];
  }
  c.request_id = r.id;
  c.mesh_id = r.mesh_id;
  c.kind = r.kind;
  c.status = status;
  c.n = r.n;
  c.slot = slot;
  c.submit_s = r.submit_s;
  c.start_s = start;
  c.done_s = now_s();
  {
    std::lock_guard<std::mutex> g(s->done_mu);
    s->done.push_back(c);
  }
  s->st.completed += 1;
  if (status != LBVH_OK) s->st.failed += 1;
}

// Marks a finished (or failed) request's mesh idle and reports it.  A
// failed build leaves no tree, and a failed refit a broken one: either way
// the mesh is no longer valid for refits.
void finish(lbvh_server *s, const Work &w, int status, uint32_t slot,
            const RequestResult *res = nullptr) {
  auto it = s->meshes.find(w.req.mesh_id);
  if (it != s->meshes.end()) {
    it->second.busy = false;
    if (w.req.kind == LBVH_REQ_BUILD) it->second.valid = (status == LBVH_OK);
    else if (status != LBVH_OK) it->second.valid = false;
  }
  complete(s, w.req, status, slot, w.start_s, res);
}

// Under the lock: gives request r its mesh range (make_room has found room
// for a new one) and marks the mesh busy.
int prepare(lbvh_server *s, Work &w) {
  Request &r = w.req;
  Mesh &m = s->meshes[r.mesh_id];
  if (r.kind == LBVH_REQ_BUILD && (m.buf == nullptr || m.n != r.n)) {
    free_mesh(s, m);
    const MeshArena::Range g = s->arena.alloc(kMeshBytesPerTri * r.n);
    if (!g.p) return LBVH_ERR_OUT_OF_MEMORY;
    m.buf = g.p;
    m.chunk = g.chunk;
    m.bytes = g.bytes;
    m.n = r.n;
    s->resident_bytes += g.bytes;
  }
  if (r.kind == LBVH_REQ_BUILD) m.valid = false;
  m.busy = true;
  m.last_use = ++s->tick;
  w.mesh = m;
  return LBVH_OK;
}

// Without the lock (slot worker): the request's device work on the slot's
// stream, then its completion event.
int enqueue(Slot &sl, Work &w) {
  const Request &r = w.req;
  const Mesh &m = w.mesh;
  hipStream_t stream = static_cast<hipStream_t>(lbvh_stream(sl.ctx));
  switch (r.source) {
    case LBVH_SRC_HOST:
      if (hipMemcpyAsync(m.tri(), r.host.data(), 48ull * r.n,
                         hipMemcpyHostToDevice, stream) != hipSuccess)
        return LBVH_ERR_DEVICE;
      break;
    case LBVH_SRC_DEVICE:
      if (hipMemcpyAsync(m.tri(), r.dev, 48ull * r.n, hipMemcpyDeviceToDevice,
                         stream) != hipSuccess)
        return LBVH_ERR_DEVICE;
      break;
    case LBVH_SRC_UNIFORM:
      lbvh::launch_synth_uniform(r.seed, r.n, r.param, m.tri(), stream);
      break;
    case LBVH_SRC_JITTER:
      lbvh::launch_synth_jitter(r.seed, r.n, r.param, m.tri(), stream);
      break;
  }
  int st;
  if (r.kind == LBVH_REQ_BUILD)
    st = lbvh_build(sl.ctx, m.tri(), r.n, nullptr, m.bvh(), m.otri(),
                    m.keys(), m.perm());
  else
    st = lbvh_refit(sl.ctx, m.tri(), r.n, m.keys(), m.perm(), m.bvh(),
                    m.otri());
  if (st != LBVH_OK) return st;
  if (w.digest) {
    const uint64_t vecs = 4ull * r.n;
    const uint32_t grid =
        (uint32_t)std::min<uint64_t>(1024, (vecs + kDigestThreads - 1) / kDigestThreads);
    hipLaunchKernelGGL(k_digest, dim3(grid), dim3(kDigestThreads), 0, stream,
                       static_cast<const uint4 *>(m.bvh()),
                       reinterpret_cast<const uint4 *>(m.otri()), r.n, sl.sums);
  }
  hipLaunchKernelGGL(k_request_done, dim3(1), dim3(1), 0, stream,
                     lbvh_context_ctl(sl.ctx), sl.sums,
                     sl.results_dev + w.ring);
  if (hipGetLastError() != hipSuccess) return LBVH_ERR_DEVICE;
  if (hipEventRecord(w.ev, stream) != hipSuccess) return LBVH_ERR_DEVICE;
  return LBVH_OK;
}

// Slot worker: takes assigned requests in order and enqueues their device
// work, so the slots' host-side launch costs overlap each other and the
// dispatcher.  It reports each enqueue's status in the Work; the dispatcher
// completes it (a failed one without waiting on its event).
void worker(lbvh_server *s, uint32_t k) {
  (void)hipSetDevice(s->device);
  Slot &sl = s->slots[k];
  std::unique_lock<std::mutex> lock(sl.m);
  while (true) {
    if (sl.mailbox.empty() && !s->stop) {
      // A blocked thread takes tens of microseconds to wake: poll for a
      // while first, while requests keep coming.
      lock.unlock();
      const double until = now_s() + kWorkerPollS;
      while (sl.mail_n.load(std::memory_order_acquire) == 0 && !s->stop &&
             now_s() < until)
        std::this_thread::yield();
      lock.lock();
    }
    sl.wcv.wait(lock, [&] { return s->stop || !sl.mailbox.empty(); });
    if (s->stop) break;
    sl.inflight.push_back(std::move(sl.mailbox.front()));
    sl.mailbox.pop_front();
    sl.mail_n.fetch_sub(1, std::memory_order_relaxed);
    Work &w = sl.inflight.back();  // deque: stays valid while others change
    lock.unlock();
    const int st = enqueue(sl, w);
    sl.scratch_bytes = lbvh_bytes_in_use(sl.ctx);
    lock.lock();
    w.status = st;
    w.launched = true;
  }
}

// Validates a queued request against the resident state.  LBVH_OK if it can
// run, another status if it must fail.
int check_request(lbvh_server *s, const Request &r) {
  const uint64_t fixed = kSlotScratchFixed * s->slots.size();
  if (s->budget &&
      kMeshBytesPerTri * r.n + kBuildScratchPerTri * r.n + fixed > s->budget)
    return LBVH_ERR_OUT_OF_BUDGET;
  if (r.kind == LBVH_REQ_REFIT || r.source == LBVH_SRC_JITTER) {
    auto it = s->meshes.find(r.mesh_id);
    if (it == s->meshes.end() || !it->second.valid) return LBVH_ERR_NOT_FOUND;
    if (it->second.n != r.n) return LBVH_ERR_SIZE_MISMATCH;
  }
  return LBVH_OK;
}

// Picks the slot for the next admission: the least loaded one with room.
uint32_t pick_slot(lbvh_server *s) {
  uint32_t best = UINT32_MAX;
  size_t best_depth = kSlotDepth;
  for (uint32_t k = 0; k < s->slots.size(); ++k) {
    const size_t d = s->slots[k].depth();
    if (d < best_depth) {
      best_depth = d;
      best = k;
    }
  }
  return best;
}

void dispatcher(lbvh_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  while (!s->stop) {
    bool progress = false;
    {
      std::lock_guard<std::mutex> g(s->in_mu);
      const uint32_t moved = (uint32_t)s->inbox.size();
      for (auto &r : s->inbox) s->queue.push_back(std::move(r));
      s->inbox.clear();
      s->inbox_n -= moved;
      s->st.submitted += moved;
    }
    // Completions, in each slot's stream order.
    for (uint32_t k = 0; k < s->slots.size(); ++k) {
      Slot &sl = s->slots[k];
      std::lock_guard<std::mutex> g(sl.m);
      while (!sl.inflight.empty() && sl.inflight.front().launched) {
        Work &w = sl.inflight.front();
        int st = w.status;
        const RequestResult *res = nullptr;
        if (st == LBVH_OK) {
          const hipError_t q = hipEventQuery(w.ev);
          if (q == hipErrorNotReady) break;
          st = q == hipSuccess ? LBVH_OK : LBVH_ERR_DEVICE;
          // the request's epilogue has written its entry by now: a tree
          // kernel that found inconsistent input fails this request only
          if (st == LBVH_OK) {
            res = &sl.results[w.ring];
            if (res->fault) st = LBVH_ERR_DEVICE;
          }
        }
        finish(s, w, st, k, res);
        sl.inflight.pop_front();
        progress = true;
      }
    }
    // Queueing deadlines, at the head of the queue whether or not a slot is
    // free (with one maximum wait they pass in arrival order).
    {
      const double now = now_s();
      size_t looked = 0;
      for (auto qi = s->queue.begin();
           qi != s->queue.end() && looked < kAdmitWindow; ++looked) {
        if (qi->deadline_s > 0.0 && now > qi->deadline_s) {
          complete(s, *qi, LBVH_ERR_TIMEOUT, UINT32_MAX, now);
          qi = s->queue.erase(qi);
          progress = true;
        } else {
          ++qi;
        }
      }
    }
    // Admissions, in arrival order.  A mesh already in a slot holds back
    // later requests on it; a request that does not fit yet holds back
    // later requests that need memory, so large builds are not starved.
    bool memory_blocked = false;
    std::unordered_set<uint64_t> held;  // meshes with an earlier request queued
    size_t scanned = 0;
    for (auto qi = s->queue.begin();
         qi != s->queue.end() && scanned < kAdmitWindow; ++scanned) {
      const uint32_t k = pick_slot(s);
      if (k == UINT32_MAX) break;
      auto mi = s->meshes.find(qi->mesh_id);
      if ((mi != s->meshes.end() && mi->second.busy) ||
          held.count(qi->mesh_id)) {
        held.insert(qi->mesh_id);
        ++qi;
        continue;
      }
      int st = check_request(s, *qi);
      if (st == LBVH_OK) {
        Slot &sl = s->slots[k];
        Need need = request_need(s, *qi, sl);
        if (need.any() && (memory_blocked || !make_room(s, *qi, sl, &need))) {
          memory_blocked = true;
          held.insert(qi->mesh_id);
          ++qi;
          continue;
        }
        sl.bound = need.slot_total;
        sl.max_build_n = need.max_build_n;
        sl.max_refit_n = need.max_refit_n;
        Work w;
        w.req = std::move(*qi);
        qi = s->queue.erase(qi);
        w.start_s = now_s();
        w.ring = (int)sl.next_event;
        w.ev = sl.events[sl.next_event];
        sl.next_event = (sl.next_event + 1) % kSlotDepth;
        w.digest = s->digest;
        st = prepare(s, w);
        if (st == LBVH_OK) {
          {
            std::lock_guard<std::mutex> g(sl.m);
            sl.mailbox.push_back(std::move(w));
            sl.mail_n.fetch_add(1, std::memory_order_release);
          }
          sl.wcv.notify_one();
          s->peak = std::max(s->peak, in_use(s));
        } else {
          finish(s, w, st, k);
        }
      } else {
        complete(s, *qi, st, UINT32_MAX, now_s());
        qi = s->queue.erase(qi);
      }
      progress = true;
    }
    const bool idle = all_idle(s);
    if (idle && s->queue.empty() && s->inbox_n == 0) s->idle_cv.notify_all();
    if (progress) {
      // let the workers and drain in before the next round
      lock.unlock();
      std::this_thread::yield();
      lock.lock();
      continue;
    }
    // While device work is in flight, poll: a timed wait sleeps for at least
    // the kernel's timer slack (about 50 us), which would be added to every
    // request's latency.  When idle, submit notifies; the timeout bounds a
    // wake-up it might race past.
    if (!idle) {
      lock.unlock();
      std::this_thread::yield();
      lock.lock();
    } else if (s->inbox_n == 0) {
      s->cv.wait_for(lock, std::chrono::microseconds(50));
    }
  }
}

}  // namespace

extern "C" {

int lbvh_server_create(int device_ordinal, uint64_t hbm_budget_bytes,
                       uint32_t slots, lbvh_server **out) {
  if (!out || slots == 0 || slots > 64) return LBVH_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  lbvh_server *s = new lbvh_server();
  s->device = device_ordinal;
  s->budget = hbm_budget_bytes;
  for (uint32_t k = 0; k < slots; ++k) s->slots.emplace_back();
  int st = LBVH_OK;
  for (auto &sl : s->slots) {
    st = lbvh_create(device_ordinal, 0, &sl.ctx);
    if (st == LBVH_OK) sl.scratch_bytes = lbvh_bytes_in_use(sl.ctx);
    for (int e = 0; e < kSlotDepth && st == LBVH_OK; ++e)
      if (hipEventCreateWithFlags(&sl.events[e], hipEventDisableTiming) !=
          hipSuccess)
        st = LBVH_ERR_DEVICE;
    if (st == LBVH_OK &&
        (hipHostMalloc((void **)&sl.results, sizeof(RequestResult) * kSlotDepth,
                       hipHostMallocCoherent) != hipSuccess ||
         hipHostGetDevicePointer((void **)&sl.results_dev, sl.results, 0) !=
             hipSuccess ||
         hipMalloc((void **)&sl.sums, 2 * sizeof(unsigned long long)) !=
             hipSuccess ||
         hipMemset(sl.sums, 0, 2 * sizeof(unsigned long long)) != hipSuccess ||
         hipDeviceSynchronize() != hipSuccess))
      st = LBVH_ERR_DEVICE;
    if (st != LBVH_OK) break;
  }
  if (st != LBVH_OK) {
    (void)hipGetLastError();
    for (auto &x : s->slots) {
      for (hipEvent_t e : x.events)
        if (e) (void)hipEventDestroy(e);
      if (x.results) (void)hipHostFree(x.results);
      if (x.sums) (void)hipFree(x.sums);
      if (x.ctx) lbvh_destroy(x.ctx);
    }
    delete s;
    return st;
  }
  for (uint32_t k = 0; k < slots; ++k)
    s->slots[k].worker = std::thread(worker, s, k);
  s->thr = std::thread(dispatcher, s);
  *out = s;
  return LBVH_OK;
}

int lbvh_server_destroy(lbvh_server *s) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  for (auto &sl : s->slots) {
    std::lock_guard<std::mutex> g(sl.m);  // no wake-up lost to the check
    sl.wcv.notify_all();
  }
  s->cv.notify_all();
  s->thr.join();
  for (auto &sl : s->slots) sl.worker.join();
  for (auto &sl : s->slots) {
    lbvh_synchronize(sl.ctx);
    for (hipEvent_t e : sl.events) (void)hipEventDestroy(e);
    (void)hipHostFree(sl.results);
    (void)hipFree(sl.sums);
  }
  (void)hipSetDevice(s->device);
  for (auto &kv : s->meshes) free_mesh(s, kv.second);
  s->arena.release_all();
  for (auto &sl : s->slots) lbvh_destroy(sl.ctx);
  delete s;
  return LBVH_OK;
}

int lbvh_server_set_div_mode(lbvh_server *s, int div_mode) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  for (auto &sl : s->slots) {
    const int st = lbvh_set_div_mode(sl.ctx, div_mode);
    if (st != LBVH_OK) return st;
  }
  s->div_mode = div_mode;
  return LBVH_OK;
}

int lbvh_server_set_digest(lbvh_server *s, int enable) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  s->digest = enable != 0;
  return LBVH_OK;
}

int lbvh_server_submit(lbvh_server *s, int kind, uint64_t mesh_id, int source,
                       const float *tri, uint32_t n, uint64_t seed,
                       double param, uint64_t *out_request_id) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  if (kind != LBVH_REQ_BUILD && kind != LBVH_REQ_REFIT) {
    std::lock_guard<std::mutex> lock(s->mu);
    s->last_error = "unknown request kind";
    return LBVH_ERR_INVALID_ARGUMENT;
  }
  if (source < LBVH_SRC_HOST || source > LBVH_SRC_JITTER ||
      (source == LBVH_SRC_JITTER && kind != LBVH_REQ_REFIT) ||
      ((source == LBVH_SRC_HOST || source == LBVH_SRC_DEVICE) && !tri) ||
      n == 0 || n > LBVH_MAX_TRIS) {
    std::lock_guard<std::mutex> lock(s->mu);
    s->last_error = "invalid source, pointer or triangle count";
    return LBVH_ERR_INVALID_ARGUMENT;
  }
  Request r;
  r.mesh_id = mesh_id;
  r.kind = kind;
  r.source = source;
  r.n = n;
  r.seed = seed;
  r.param = param;
  r.submit_s = now_s();
  const double wait = s->max_wait.load();
  if (wait > 0.0) r.deadline_s = r.submit_s + wait;
  if (source == LBVH_SRC_HOST) r.host.assign(tri, tri + 12ull * n);
  if (source == LBVH_SRC_DEVICE) r.dev = tri;
  {
    std::lock_guard<std::mutex> lock(s->in_mu);
    r.id = s->next_id++;
    if (out_request_id) *out_request_id = r.id;
    s->inbox.push_back(std::move(r));
    s->inbox_n += 1;
  }
  s->cv.notify_all();
  return LBVH_OK;
}

int lbvh_server_set_max_wait(lbvh_server *s, double max_wait_s) {
  if (!s || !(max_wait_s >= 0.0)) return LBVH_ERR_INVALID_ARGUMENT;
  s->max_wait = max_wait_s;
  return LBVH_OK;
}

int lbvh_server_cancel(lbvh_server *s, uint64_t request_id) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  bool found = false;
  Request r;
  {
    std::lock_guard<std::mutex> g(s->in_mu);
    for (auto it = s->inbox.begin(); it != s->inbox.end(); ++it)
      if (it->id == request_id) {
        r = std::move(*it);
        s->inbox.erase(it);
        s->inbox_n -= 1;
        s->st.submitted += 1;  // as if the dispatcher had moved it
        found = true;
        break;
      }
  }
  if (!found)
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
      if (it->id == request_id) {
        r = std::move(*it);
        s->queue.erase(it);
        found = true;
        break;
      }
  if (found) {
    complete(s, r, LBVH_ERR_CANCELLED, UINT32_MAX, now_s());
    s->cv.notify_all();
    return LBVH_OK;
  }
  for (auto &sl : s->slots) {
    std::lock_guard<std::mutex> g(sl.m);
    for (auto *d : {&sl.mailbox, &sl.inflight})
      for (auto &w : *d)
        if (w.req.id == request_id) {
          s->last_error = "request is already admitted";
          return LBVH_ERR_BUSY;
        }
  }
  s->last_error = "no waiting request with this id";
  return LBVH_ERR_NOT_FOUND;
}

int lbvh_server_poll(lbvh_server *s, lbvh_completion *out, uint32_t max,
                     uint32_t *out_count) {
  if (!s || !out_count || (max && !out)) return LBVH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->done_mu);
  uint32_t k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *out_count = k;
  return LBVH_OK;
}

int lbvh_server_drain(lbvh_server *s) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lock(s->mu);
  s->cv.notify_all();
  s->idle_cv.wait(lock, [s] {
    return s->queue.empty() && s->inbox_n == 0 && all_idle(s);
  });
  for (auto &sl : s->slots) {
    const int st = lbvh_synchronize(sl.ctx);  // also reports kernel faults
    if (st != LBVH_OK) {
      s->last_error = lbvh_last_error(sl.ctx);
      return st;
    }
  }
  return LBVH_OK;
}

int lbvh_server_mesh(lbvh_server *s, uint64_t mesh_id, void **bvh,
                     float **tri_out, uint32_t *n) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->meshes.find(mesh_id);
  if (it == s->meshes.end() || !it->second.valid) {
    s->last_error = "mesh is not resident";
    return LBVH_ERR_NOT_FOUND;
  }
  if (bvh) *bvh = it->second.bvh();
  if (tri_out) *tri_out = it->second.otri();
  if (n) *n = it->second.n;
  return LBVH_OK;
}

int lbvh_server_evict(lbvh_server *s, uint64_t mesh_id) {
  if (!s) return LBVH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->meshes.find(mesh_id);
  if (it == s->meshes.end()) return LBVH_ERR_NOT_FOUND;
  if (it->second.busy) return LBVH_ERR_BUSY;
  free_mesh(s, it->second);
  s->meshes.erase(it);
  s->st.evictions += 1;
  return LBVH_OK;
}

int lbvh_server_stats_get(lbvh_server *s, lbvh_server_stats *out) {
  if (!s || !out) return LBVH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  *out = s->st;
  out->submitted += s->inbox_n;  // not yet moved to the queue
  out->resident_meshes = 0;
  for (auto &kv : s->meshes) out->resident_meshes += kv.second.buf ? 1 : 0;
  out->resident_bytes = s->resident_bytes;
  out->bytes_in_use = in_use(s);
  out->peak_bytes = s->peak;
  out->queued = (uint32_t)s->queue.size() + s->inbox_n;
  uint32_t busy = 0;
  for (auto &sl : s->slots) busy += (uint32_t)sl.depth();
  out->in_flight = busy;
  return LBVH_OK;
}

const char *lbvh_server_last_error(lbvh_server *s) {
  return s ? s->last_error.c_str() : "null server";
}

}  // extern "C"
