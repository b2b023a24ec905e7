// Randomised check of the server's mesh-memory range allocator on the CPU:
// live ranges never overlap and stay inside their chunk, the allocator's
// byte accounting balances, a freed chunk merges back into one free range,
// and empty chunks are returned.  A second run makes every k-th chunk
// allocation fail (as the device allocator can) and checks that a failed
// growth changes nothing: the live ranges (resident meshes) keep their
// place and their bytes, and the allocator goes on working.  Exit status 0
// on success.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <cmath>
#include <vector>

#include "range_arena.h"

namespace {
int g_live_chunks = 0;
uint64_t g_alloc_calls = 0;
uint64_t g_fail_every = 0;  // 0: never fail
struct HostChunks {
  static void *alloc(uint64_t bytes) {
    ++g_alloc_calls;
    if (g_fail_every && g_alloc_calls % g_fail_every == 0) return nullptr;
    ++g_live_chunks;
    return std::aligned_alloc(256, bytes);
  }
  static void release(void *p) {
    --g_live_chunks;
    std::free(p);
  }
};
using Arena = lbvh::RangeArena<HostChunks>;

#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                      \
    }                                                                \
  } while (0)

int run(uint32_t seed) {
  std::mt19937_64 rng(seed);
  const uint64_t budget = 64ull << 20;  // chunks of 4 MiB
  Arena a;
  std::vector<Arena::Range> live;
  std::vector<std::pair<char *, uint64_t>> chunk_spans;
  uint64_t used = 0;
  for (int step = 0; step < 20000; ++step) {
    const bool do_alloc = live.empty() || (rng() % 100) < 55;
    if (do_alloc) {
      // log-uniform sizes 1 B .. 6 MiB (some above the chunk size)
      const double e = std::uniform_real_distribution<double>(0, 22.6)(rng);
      const uint64_t bytes = 1 + (uint64_t)std::exp2(e);
      if (!a.can_fit(bytes)) {
        const uint64_t c = Arena::chunk_for(bytes, budget);
        CHECK(c >= Arena::round(bytes));
        CHECK(a.add_chunk(c));
      }
      CHECK(a.can_fit(bytes));
      Arena::Range g = a.alloc(bytes);
      CHECK(g.p != nullptr && g.bytes == Arena::round(bytes));
      CHECK(((uintptr_t)g.p & 255) == 0);
      for (auto &o : live)
        CHECK(g.p + g.bytes <= o.p || o.p + o.bytes <= g.p);
      live.push_back(g);
      used += g.bytes;
    } else {
      const size_t k = rng() % live.size();
      a.free(live[k]);
      used -= live[k].bytes;
      live[k] = live.back();
      live.pop_back();
    }
    CHECK(a.free_bytes() + used == a.reserved());
    if (step % 997 == 0 && (rng() & 1)) a.release_empty();
    CHECK((int)a.chunks() == g_live_chunks);
  }
  for (auto &g : live) a.free(g);
  // every chunk merged back into one free range
  CHECK(a.free_ranges() == a.chunks());
  CHECK(a.free_bytes() == a.reserved());
  CHECK(a.release_empty());
  CHECK(a.reserved() == 0 && a.chunks() == 0 && g_live_chunks == 0);
  return 0;
}
// Every k-th chunk allocation fails.  Each live range carries its id in its
// first and last 8 bytes; a failed growth must leave the allocator's state
// and every live range's bytes as they were.  On a failure the caller does
// what the server's admission does: frees a resident range and tries again,
// or gives the request up when nothing is left to free.
int run_with_failures(uint32_t seed, uint64_t every) {
  std::mt19937_64 rng(seed);
  g_fail_every = every;
  g_alloc_calls = 0;
  const uint64_t budget = 64ull << 20;
  Arena a;
  struct Live {
    Arena::Range g;
    uint64_t id;
  };
  std::vector<Live> live;
  uint64_t next_id = 1, failures = 0, gave_up = 0;
  const auto tag = [](const Arena::Range &g, uint64_t id) {
    std::memcpy(g.p, &id, 8);
    std::memcpy(g.p + g.bytes - 8, &id, 8);
  };
  const auto intact = [&]() {
    for (auto &l : live) {
      uint64_t x = 0, y = 0;
      std::memcpy(&x, l.g.p, 8);
      std::memcpy(&y, l.g.p + l.g.bytes - 8, 8);
      if (x != l.id || y != l.id) return false;
    }
    return true;
  };
  for (int step = 0; step < 6000; ++step) {
    if (live.empty() || (rng() % 100) < 55) {
      const double e = std::uniform_real_distribution<double>(3, 22.6)(rng);
      const uint64_t bytes = 16 + (uint64_t)std::exp2(e);
      bool placed = false;
      while (!placed) {
        if (a.can_fit(bytes)) {
          Arena::Range g = a.alloc(bytes);
          CHECK(g.p != nullptr);
          tag(g, next_id);
          live.push_back({g, next_id++});
          placed = true;
          break;
        }
        const uint64_t reserved = a.reserved();
        const size_t chunks = a.chunks(), ranges = a.free_ranges();
        const uint64_t freeb = a.free_bytes();
        if (a.add_chunk(Arena::chunk_for(bytes, budget))) continue;
        ++failures;
        // nothing changed, and no resident range moved or was overwritten
        CHECK(a.reserved() == reserved && a.chunks() == chunks);
        CHECK(a.free_ranges() == ranges && a.free_bytes() == freeb);
        CHECK(intact());
        if (live.empty()) {
          ++gave_up;
          break;
        }
        const size_t k = rng() % live.size();  // evict one, try again
        a.free(live[k].g);
        live[k] = live.back();
        live.pop_back();
        a.release_empty();
      }
    } else {
      const size_t k = rng() % live.size();
      a.free(live[k].g);
      live[k] = live.back();
      live.pop_back();
    }
    CHECK(intact());
    CHECK((int)a.chunks() == g_live_chunks);
  }
  CHECK(failures > 0);
  for (auto &l : live) a.free(l.g);
  CHECK(a.free_bytes() == a.reserved());
  a.release_empty();
  CHECK(a.reserved() == 0 && g_live_chunks == 0);
  g_fail_every = 0;
  (void)gave_up;
  return 0;
}
}  // namespace

int main() {
  for (uint32_t seed = 1; seed <= 8; ++seed)
    if (run(seed)) {
      std::fprintf(stderr, "seed %u\n", seed);
      return 1;
    }
  for (uint32_t seed = 1; seed <= 6; ++seed)
    for (uint64_t every : {1ull, 2ull, 3ull, 7ull})
      if (run_with_failures(seed, every)) {
        std::fprintf(stderr, "seed %u, every %llu\n", seed,
                     (unsigned long long)every);
        return 1;
      }
  std::puts("ok");
  return 0;
}
