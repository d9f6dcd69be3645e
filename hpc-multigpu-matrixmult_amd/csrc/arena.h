// First-fit offset allocator with coalescing free blocks: the bookkeeping of the serving
// loop's operand arena (one device allocation sub-allocated at admission). Host-only,
// no HIP dependency.
#ifndef PHPC_ARENA_H
#define PHPC_ARENA_H

#include <cstddef>
#include <iterator>
#include <map>

namespace phpc {

class OffsetArena {
 public:
  static constexpr size_t kAlign = 256;
  static constexpr size_t kNone = ~(size_t)0;
  static size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

  // capacity rounded down to the alignment; returns it
  size_t reset(size_t bytes) {
    cap_ = bytes / kAlign * kAlign;
    free_.clear();
    if (cap_) free_[0] = cap_;
    return cap_;
  }
  size_t capacity() const { return cap_; }
  size_t free_blocks() const { return free_.size(); }

  // offset of a block of at least `bytes` (rounded up), or kNone
  size_t alloc(size_t bytes) {
    bytes = round_up(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      return off;
    }
    return kNone;
  }
  // give back a block from alloc (same size argument); neighbours merge
  void release(size_t off, size_t bytes) {
    bytes = round_up(bytes ? bytes : 1);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  size_t cap_ = 0;
  std::map<size_t, size_t> free_;  // offset -> length
};

}  // namespace phpc

#endif
