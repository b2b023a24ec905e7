// range_arena.h -- best-fit range allocator over large chunks (host code).
//
// The server's mesh memory: chunks come from Chunks::alloc (device memory in
// the server, plain memory in the allocator's CPU test) and are carved into
// 256-byte-aligned ranges, smallest fitting free range first, with a freed
// range merged into free neighbours of the same chunk.  Not thread-safe: the
// server calls it under its lock.
#pragma once

#include <algorithm>
#include <cstdint>
#include <map>
#include <vector>

namespace lbvh {

template <class Chunks>
class RangeArena {
 public:
  static constexpr uint64_t kChunkBytes = 4ull << 30;
  static constexpr uint64_t kAlign = 256;

  struct Range {
    char *p = nullptr;
    uint32_t chunk = 0;
    uint64_t bytes = 0;
  };

  static uint64_t round(uint64_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }
  // Size of a new chunk for a range of `bytes`: 4 GiB (at most 1/16 of a
  // nonzero budget, so a small budget is not taken by one chunk), or the
  // range itself when larger.
  static uint64_t chunk_for(uint64_t bytes, uint64_t budget) {
    const uint64_t c = budget ? std::min(kChunkBytes, budget / 16) : kChunkBytes;
    return std::max(round(c), round(bytes));
  }

  ~RangeArena() { release_all(); }

  uint64_t reserved() const { return reserved_; }
  size_t chunks() const {
    size_t k = 0;
    for (auto &c : chunks_) k += c.p != nullptr;
    return k;
  }
  bool can_fit(uint64_t bytes) const {
    return by_size_.lower_bound(round(bytes)) != by_size_.end();
  }
  bool add_chunk(uint64_t bytes) {
    bytes = round(bytes);
    char *p = static_cast<char *>(Chunks::alloc(bytes));
    if (!p) return false;
    uint32_t id = 0;
    while (id < chunks_.size() && chunks_[id].p) ++id;
    if (id == chunks_.size()) chunks_.emplace_back();
    chunks_[id] = {p, bytes, 0};
    reserved_ += bytes;
    insert_free(p, bytes, id);
    return true;
  }
  // The smallest free range that holds `bytes`; its tail stays free.  An
  // empty Range if none does.
  Range alloc(uint64_t bytes) {
    bytes = round(bytes);
    auto it = by_size_.lower_bound(bytes);
    if (it == by_size_.end()) return {};
    char *p = reinterpret_cast<char *>(it->second);
    by_size_.erase(it);
    auto f = by_addr_.find(key(p));
    const Free r = f->second;
    by_addr_.erase(f);
    if (r.bytes > bytes) insert_free(p + bytes, r.bytes - bytes, r.chunk);
    chunks_[r.chunk].used += bytes;
    return {p, r.chunk, bytes};
  }
  void free(const Range &g) {
    chunks_[g.chunk].used -= g.bytes;
    char *p = g.p;
    uint64_t n = g.bytes;
    auto nx = by_addr_.find(key(p + n));
    if (nx != by_addr_.end() && nx->second.chunk == g.chunk) {
      n += nx->second.bytes;
      erase_size(nx->first, nx->second.bytes);
      by_addr_.erase(nx);
    }
    auto pv = by_addr_.lower_bound(key(p));
    if (pv != by_addr_.begin()) {
      --pv;
      if (pv->second.chunk == g.chunk &&
          pv->first + pv->second.bytes == key(p)) {
        p = reinterpret_cast<char *>(pv->first);
        n += pv->second.bytes;
        erase_size(pv->first, pv->second.bytes);
        by_addr_.erase(pv);
      }
    }
    insert_free(p, n, g.chunk);
  }
  // Returns the chunks that hold no range to Chunks::release; true if there
  // were any.
  bool release_empty() {
    bool any = false;
    for (auto &c : chunks_) {
      if (!c.p || c.used) continue;
      auto f = by_addr_.find(key(c.p));
      erase_size(f->first, f->second.bytes);
      by_addr_.erase(f);
      Chunks::release(c.p);
      reserved_ -= c.bytes;
      c = Chunk{};
      any = true;
    }
    return any;
  }
  void release_all() {
    for (auto &c : chunks_)
      if (c.p) Chunks::release(c.p);
    chunks_.clear();
    by_addr_.clear();
    by_size_.clear();
    reserved_ = 0;
  }
  // Free bytes in ranges (for checks).
  uint64_t free_bytes() const {
    uint64_t b = 0;
    for (auto &kv : by_addr_) b += kv.second.bytes;
    return b;
  }
  size_t free_ranges() const { return by_addr_.size(); }

 private:
  struct Free {
    uint64_t bytes;
    uint32_t chunk;
  };
  struct Chunk {
    char *p = nullptr;
    uint64_t bytes = 0, used = 0;
  };
  static uintptr_t key(const char *p) { return reinterpret_cast<uintptr_t>(p); }
  void insert_free(char *p, uint64_t n, uint32_t c) {
    by_addr_[key(p)] = {n, c};
    by_size_.emplace(n, key(p));
  }
  void erase_size(uintptr_t p, uint64_t n) {
    auto r = by_size_.equal_range(n);
    for (auto i = r.first; i != r.second; ++i)
      if (i->second == p) {
        by_size_.erase(i);
        return;
      }
  }
  std::vector<Chunk> chunks_;
  std::map<uintptr_t, Free> by_addr_;           // free ranges by address
  std::multimap<uint64_t, uintptr_t> by_size_;  // the same, by size
  uint64_t reserved_ = 0;
};

}  // namespace lbvh
