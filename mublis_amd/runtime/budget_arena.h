// Best-fit sub-allocator over one reserved address range.
//
// The serving loop reserves its whole HBM budget once and carves operand
// matrices and launch workspace out of it, so admission is a bookkeeping
// decision (does the block fit?) and a release never calls hipFree, which
// would wait for the whole device.  Offsets are relative to the range start.
// Free blocks are kept twice: by offset (for coalescing with neighbours) and
// by size (for best fit).  Not thread-safe: the owner holds its own lock.
// Pure C++ so a host-only build can exercise it under the sanitizers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>
#include <set>
#include <utility>

namespace mublis_serve {

class BudgetArena {
 public:
  static constexpr size_t kAlign = 256;
  static constexpr size_t kFail = ~size_t(0);

  explicit BudgetArena(size_t capacity = 0) { reset(capacity); }

  void reset(size_t capacity) {
    cap_ = capacity / kAlign * kAlign;
    used_ = peak_ = 0;
    by_off_.clear();
    by_size_.clear();
    live_.clear();
    if (cap_) insert_free(0, cap_);
  }

  static size_t round(size_t bytes) {
    return (bytes == 0 ? kAlign : (bytes + kAlign - 1) / kAlign * kAlign);
  }

  // Offset of a block of at least `bytes`, or kFail when no free block fits.
  size_t alloc(size_t bytes) {
    const size_t need = round(bytes);
    auto it = by_size_.lower_bound({need, 0});
    if (it == by_size_.end()) return kFail;
    const size_t off = it->second, len = it->first;
    erase_free(off, len);
    if (len > need) insert_free(off + need, len - need);
    live_[off] = need;
    used_ += need;
    if (used_ > peak_) peak_ = used_;
    return off;
  }

  // Returns false for an offset that is not a live block (double free).
  bool release(size_t off) {
    auto lv = live_.find(off);
    if (lv == live_.end()) return false;
    size_t len = lv->second;
    live_.erase(lv);
    used_ -= len;
    size_t start = off;
    auto next = by_off_.lower_bound(off);
    if (next != by_off_.end() && next->first == off + len) {
      const size_t nlen = next->second;
      erase_free(next->first, nlen);
      len += nlen;
    }
    auto prev = by_off_.lower_bound(start);
    if (prev != by_off_.begin()) {
      --prev;
      if (prev->first + prev->second == start) {
        start = prev->first;
        const size_t plen = prev->second;
        erase_free(start, plen);
        len += plen;
      }
    }
    insert_free(start, len);
    return true;
  }

  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t largest_free() const { return by_size_.empty() ? 0 : by_size_.rbegin()->first; }
  size_t free_blocks() const { return by_off_.size(); }
  size_t live_blocks() const { return live_.size(); }
  void reset_peak() { peak_ = used_; }

 private:
  void insert_free(size_t off, size_t len) {
    by_off_[off] = len;
    by_size_.insert({len, off});
  }
  void erase_free(size_t off, size_t len) {
    by_off_.erase(off);
    by_size_.erase({len, off});
  }

  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> by_off_;             // free: offset -> length
  std::set<std::pair<size_t, size_t>> by_size_;  // free: (length, offset)
  std::map<size_t, size_t> live_;               // allocated: offset -> length
};

}  // namespace mublis_serve
