// alloc_table.h -- extents of the device buffers a context handed out
// (host code).
//
// lbvh_alloc records each buffer here.  Before an entry point copies into or
// launches over a pointer, it asks whether [p, p + bytes) stays inside the
// buffer that contains p.  A pointer inside no recorded buffer (the caller's
// own hipMalloc) is not checked.  The reference never needs this check:
// WebGPU bounds-checks every buffer access, and mappedAtCreation +
// Float32Array.set throws RangeError on an oversize copy (build-bvh.js:14-21).
// Not thread-safe: the context calls it under its lock.
#pragma once

#include <cstdint>
#include <map>

namespace lbvh {

class AllocTable {
 public:
  enum Extent {
    kForeign = 0,  // p lies in no recorded buffer
    kInside = 1,   // [p, p + bytes) lies in the buffer containing p
    kOverrun = 2   // [p, p + bytes) runs past the end of that buffer
  };

  void add(const void *p, uint64_t bytes) { m_[addr(p)] = bytes; }

  // The size of the buffer starting exactly at p; false if there is none.
  bool find(const void *p, uint64_t *bytes) const {
    auto it = m_.find(addr(p));
    if (it == m_.end()) return false;
    if (bytes) *bytes = it->second;
    return true;
  }

  // Removes the buffer starting exactly at p; false if there is none.
  bool remove(const void *p, uint64_t *bytes) {
    auto it = m_.find(addr(p));
    if (it == m_.end()) return false;
    if (bytes) *bytes = it->second;
    m_.erase(it);
    return true;
  }

  bool empty() const { return m_.empty(); }
  size_t size() const { return m_.size(); }

  template <class F>
  void for_each(F f) const {
    for (auto &kv : m_) f(reinterpret_cast<void *>(kv.first), kv.second);
  }
  void clear() { m_.clear(); }

  // The buffer containing p is the one with the largest start <= p whose
  // range [start, start + size) holds p; a zero-size buffer contains only
  // its own start.  `room` receives the bytes from p to that buffer's end.
  Extent check(const void *p, uint64_t bytes, uint64_t *room = nullptr) const {
    const uintptr_t a = addr(p);
    auto it = m_.upper_bound(a);
    if (it == m_.begin()) return kForeign;
    --it;
    const uintptr_t start = it->first;
    const uint64_t size = it->second;
    const bool contains = (a == start) || (a - start < size);
    if (!contains) return kForeign;
    const uint64_t left = size - (a - start);
    if (room) *room = left;
    return bytes <= left ? kInside : kOverrun;
  }

 private:
  static uintptr_t addr(const void *p) { return reinterpret_cast<uintptr_t>(p); }
  std::map<uintptr_t, uint64_t> m_;
};

}  // namespace lbvh
