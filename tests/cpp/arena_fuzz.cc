// Randomised check of the server's mesh-memory range allocator on the CPU:
// live ranges never overlap and stay inside their chunk, the allocator's
// byte accounting balances, a freed chunk merges back into one free range,
// and empty chunks are returned.  Exit status 0 on success.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <cmath>
#include <vector>

#include "range_arena.h"

namespace {
int g_live_chunks = 0;
struct HostChunks {
  static void *alloc(uint64_t bytes) {
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
}  // namespace

int main() {
  for (uint32_t seed = 1; seed <= 8; ++seed)
    if (run(seed)) {
      std::fprintf(stderr, "seed %u\n", seed);
      return 1;
    }
  std::puts("ok");
  return 0;
}
