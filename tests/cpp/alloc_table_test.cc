// Checks the extent table behind the C ABI's buffer checks on the CPU:
// a copy or launch over [p, p + bytes) is accepted only when it stays in the
// lbvh_alloc buffer containing p, pointers in no buffer are left to the
// caller, and add/find/remove keep the table consistent.  Random buffers are
// compared with a brute-force scan.  Exit status 0 on success.
#include <cstdio>
#include <random>
#include <vector>

#include "alloc_table.h"

using lbvh::AllocTable;

#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

namespace {

struct Buf {
  uintptr_t start;
  uint64_t size;
};

// Brute force: the buffer whose [start, start + size) holds a (or whose
// start is a, for a zero-size buffer).
AllocTable::Extent brute(const std::vector<Buf> &bufs, uintptr_t a,
                         uint64_t bytes) {
  for (const Buf &b : bufs) {
    const bool in = (a == b.start) || (a > b.start && a < b.start + b.size);
    if (in) return bytes <= b.start + b.size - a ? AllocTable::kInside
                                                 : AllocTable::kOverrun;
  }
  return AllocTable::kForeign;
}

const void *P(uintptr_t a) { return reinterpret_cast<const void *>(a); }

int known_answers() {
  AllocTable t;
  t.add(P(0x1000), 0x100);  // [0x1000, 0x1100)
  t.add(P(0x1100), 0);      // zero-size buffer right after it
  t.add(P(0x2000), 48);
  CHECK(t.check(P(0x1000), 0x100) == AllocTable::kInside);
  CHECK(t.check(P(0x1000), 0x101) == AllocTable::kOverrun);
  CHECK(t.check(P(0x10ff), 1) == AllocTable::kInside);
  CHECK(t.check(P(0x10ff), 2) == AllocTable::kOverrun);
  uint64_t room = 0;
  CHECK(t.check(P(0x1080), 0x81, &room) == AllocTable::kOverrun && room == 0x80);
  CHECK(t.check(P(0x1100), 0) == AllocTable::kInside);  // the empty buffer
  CHECK(t.check(P(0x1100), 1) == AllocTable::kOverrun);
  CHECK(t.check(P(0x1101), 4) == AllocTable::kForeign);
  CHECK(t.check(P(0x0fff), 1 << 20) == AllocTable::kForeign);
  CHECK(t.check(P(0x2030), 1) == AllocTable::kForeign);  // one past the end
  CHECK(t.check(P(0x2000), 48) == AllocTable::kInside);
  CHECK(t.check(P(0x2000), 49) == AllocTable::kOverrun);
  uint64_t sz = 0;
  CHECK(t.find(P(0x2000), &sz) && sz == 48);
  CHECK(!t.find(P(0x2001), &sz));
  CHECK(!t.remove(P(0x2001), nullptr));
  CHECK(t.remove(P(0x2000), &sz) && sz == 48 && t.size() == 2);
  CHECK(t.check(P(0x2000), 1) == AllocTable::kForeign);
  return 0;
}

int random_layouts(uint32_t seed) {
  std::mt19937_64 rng(seed);
  AllocTable t;
  std::vector<Buf> bufs;
  uintptr_t at = 0x10000;
  for (int i = 0; i < 300; ++i) {
    at += 256 * (rng() % 4);  // gaps of 0..768 bytes, sometimes adjacent
    const uint64_t size = (rng() % 8 == 0) ? 0 : 1 + rng() % 5000;
    bufs.push_back({at, size});
    t.add(P(at), size);
    at += (size + 255) & ~uint64_t(255);
    if (size == 0) at += 256;
  }
  for (int q = 0; q < 200000; ++q) {
    const uintptr_t a = 0x10000 - 512 + rng() % (at - 0x10000 + 1024);
    const uint64_t bytes = (rng() % 4 == 0) ? 0 : rng() % 6000;
    CHECK(t.check(P(a), bytes) == brute(bufs, a, bytes));
  }
  // remove half and check again
  std::vector<Buf> kept;
  for (const Buf &b : bufs) {
    if (rng() % 2) {
      uint64_t sz = 0;
      CHECK(t.remove(P(b.start), &sz) && sz == b.size);
    } else {
      kept.push_back(b);
    }
  }
  CHECK(t.size() == kept.size());
  for (int q = 0; q < 100000; ++q) {
    const uintptr_t a = 0x10000 - 512 + rng() % (at - 0x10000 + 1024);
    const uint64_t bytes = rng() % 6000;
    CHECK(t.check(P(a), bytes) == brute(kept, a, bytes));
  }
  return 0;
}

}  // namespace

int main() {
  if (known_answers()) return 1;
  for (uint32_t seed = 1; seed <= 5; ++seed)
    if (random_layouts(seed)) return 1;
  std::puts("ok");
  return 0;
}
