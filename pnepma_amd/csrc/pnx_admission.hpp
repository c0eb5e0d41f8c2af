// Host-side pieces of the server's admission that need no device: the
// first-fit allocator of the HBM arena and the taped/checkpointed/wait
// decision.  Header-only, so that a host test can build them alone.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace pnx {

constexpr size_t kAlign = 256;

inline size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
// First-fit allocator of offsets within a range of `size` bytes.
class FreeList {
 public:
  explicit FreeList(size_t size = 0) : size_(size) {
    if (size) free_[0] = size;
  }
  // returns the offset, or SIZE_MAX when no free range is large enough
  size_t alloc(size_t bytes) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_ += bytes;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    used_ -= bytes;
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
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

// How an arrived request is admitted.
enum Admission : int { kAdmitTaped = 0, kAdmitCheckpointed = 1, kAdmitWait = 2 };

// Taped (its forward states kept, no recompute) while the arena's bytes in
// use plus its taped size stay within the tape limit; else checkpointed if its
// lane is short of work (fewer than busy_limit active requests) or taping is
// off (tape_limit 0); else it waits for memory.  Whether the chosen layout
// fits in the arena is the allocator's answer.
inline Admission admission(size_t used, size_t tape_limit, size_t bytes_tape, int lane_active, int busy_limit) {
  if (used + bytes_tape <= tape_limit) return kAdmitTaped;
  if (tape_limit > 0 && lane_active >= busy_limit) return kAdmitWait;
  return kAdmitCheckpointed;
}

}  // namespace pnx
