// pool.hpp -- best-fit allocator over one address range (host logic only;
// the serving loop places it over a single device allocation).
#pragma once
#include <algorithm>
#include <cstdint>
#include <map>
#include <set>
#include <unordered_map>
#include <utility>

namespace par2hip {

constexpr uint64_t kGranule = 256;  // block sizes and offsets are multiples of this

// Free blocks are kept by offset (for coalescing) and by (length, offset)
// (for best fit: the shortest block that holds the request, lowest offset
// among equals).  Not thread-safe.
class RangePool {
 public:
  void init(uint64_t bytes) {
    by_off_.clear();
    by_size_.clear();
    used_.clear();
    size_ = bytes & ~(kGranule - 1);
    if (size_) insert_free(0, size_);
  }
  // offset of a block of at least `bytes`, or UINT64_MAX when none fits
  uint64_t alloc(uint64_t bytes) {
    bytes = std::max<uint64_t>((bytes + kGranule - 1) & ~(kGranule - 1), kGranule);
    auto it = by_size_.lower_bound({bytes, 0});
    if (it == by_size_.end()) return UINT64_MAX;
    const uint64_t len = it->first, off = it->second;
    erase_free(off, len);
    if (len > bytes) insert_free(off + bytes, len - bytes);
    used_[off] = bytes;
    in_use_ += bytes;
    return off;
  }
  // returns false for an offset that is not an allocated block
  bool release(uint64_t off) {
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    uint64_t start = off, len = u->second;
    in_use_ -= len;
    used_.erase(u);
    auto next = by_off_.find(start + len);
    if (next != by_off_.end()) {
      const uint64_t nl = next->second;
      erase_free(start + len, nl);
      len += nl;
    }
    auto prev = by_off_.lower_bound(start);
    if (prev != by_off_.begin()) {
      --prev;
      if (prev->first + prev->second == start) {
        const uint64_t po = prev->first, pl = prev->second;
        erase_free(po, pl);
        start = po;
        len += pl;
      }
    }
    insert_free(start, len);
    return true;
  }
  uint64_t size() const { return size_; }
  uint64_t in_use() const { return in_use_; }
  size_t free_blocks() const { return by_off_.size(); }
  uint64_t largest_free() const { return by_size_.empty() ? 0 : by_size_.rbegin()->first; }

 private:
  void insert_free(uint64_t off, uint64_t len) {
    by_off_[off] = len;
    by_size_.insert({len, off});
  }
  void erase_free(uint64_t off, uint64_t len) {
    by_off_.erase(off);
    by_size_.erase({len, off});
  }
  uint64_t size_ = 0, in_use_ = 0;
  std::map<uint64_t, uint64_t> by_off_;              // free: offset -> length
  std::set<std::pair<uint64_t, uint64_t>> by_size_;  // free: (length, offset)
  std::unordered_map<uint64_t, uint64_t> used_;      // allocated: offset -> length
};

}  // namespace par2hip
