// Host-side bookkeeping of the serving runtime (no device calls, so it is
// unit-tested on the CPU):
//   Arena      first-fit allocator of offsets inside one device range,
//              256-byte blocks, freed neighbours coalesce;
//   Admission  strict-FIFO admission of requests under a byte budget.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>

namespace gfft {

class Arena {
 public:
  static constexpr size_t kAlign = 256;
  static constexpr size_t kNone = SIZE_MAX;

  explicit Arena(size_t bytes = 0) { reset(bytes); }

  void reset(size_t bytes) {
    free_.clear();
    size_ = bytes / kAlign * kAlign;
    used_ = 0;
    if (size_) free_[0] = size_;
  }

  static size_t round(size_t n) { return (std::max<size_t>(n, 1) + kAlign - 1) / kAlign * kAlign; }

  // Offset of a block of at least n bytes (lowest address that fits), or kNone.
  size_t alloc(size_t n) {
    n = round(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const size_t off = it->first, rest = it->second - n;
      free_.erase(it);
      if (rest) free_[off + n] = rest;
      used_ += n;
      return off;
    }
    return kNone;
  }

  // Returns a block obtained from alloc(n) with the same n.
  void release(size_t off, size_t n) {
    n = round(n);
    used_ -= n;
    auto it = free_.emplace(off, n).first;
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

  // A block of n bytes can exist at all (an empty arena would have one).
  bool can_hold(size_t n) const { return round(n) <= size_; }

  size_t size() const { return size_; }
  size_t used() const { return used_; }
  size_t free_blocks() const { return free_.size(); }
  size_t largest_free() const {
    size_t m = 0;
    for (const auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }

 private:
  std::map<size_t, size_t> free_;  // offset -> length, disjoint, never adjacent
  size_t size_ = 0, used_ = 0;
};

// Requests wait in arrival order in one of kQueues FIFOs (the server keeps
// requests at or under its small-lane threshold in queue 1, larger ones in
// queue 0); a FIFO's head is admitted once its bytes fit under the budget
// next to everything in flight (and, for requests staged through the arena,
// once a block is free).  A request never overtakes an earlier one of its
// own FIFO, so a large request cannot starve; the FIFOs share the budget.
template <typename Item>
class Admission {
 public:
  static constexpr int kQueues = 2;
  explicit Admission(size_t budget = 0) : budget_(budget) {}

  size_t budget() const { return budget_; }
  size_t in_flight() const { return in_flight_; }
  size_t peak() const { return peak_; }
  size_t waiting() const { return q_[0].size() + q_[1].size(); }
  bool fits(size_t n) const { return n <= budget_ && in_flight_ + n <= budget_; }

  void push(const Item& it, int q = 0) { q_[q].push_back(it); }
  bool empty() const { return q_[0].empty() && q_[1].empty(); }
  bool empty(int q) const { return q_[q].empty(); }
  Item& head(int q = 0) { return q_[q].front(); }
  void pop_admitted(size_t n, int q = 0) {
    q_[q].pop_front();
    in_flight_ += n;
    peak_ = std::max(peak_, in_flight_);
  }
  void finished(size_t n) { in_flight_ -= n; }

 private:
  std::deque<Item> q_[kQueues];
  size_t budget_ = 0, in_flight_ = 0, peak_ = 0;
};

// Stream choice at admission.  depth[i] / bytes[i]: unfinished requests and
// their array bytes on stream i.  With a reserved stream (reserve: the last
// one), requests that are not small never use it.  A stream is eligible
// while it holds fewer than max_depth requests (0 = no limit); among the
// eligible ones the one with the least outstanding bytes wins (lowest index
// on ties).  Returns -1 when no stream is eligible: the request keeps
// waiting in its FIFO.
inline int pick_stream(const size_t* depth, const size_t* bytes, int n, bool reserve, bool small, size_t max_depth) {
  int best = -1;
  for (int i = 0; i < n; ++i) {
    if (reserve && !small && i == n - 1) continue;
    if (max_depth && depth[i] >= max_depth) continue;
    if (best < 0 || bytes[i] < bytes[best]) best = i;
  }
  return best;
}

}  // namespace gfft
