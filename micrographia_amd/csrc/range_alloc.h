// First-fit range allocator of the batched engine's arenas (slots, list entries, cells): ranges
// of [0, capacity) in units, the lowest free range that fits is taken (starting at a multiple of
// `align` when one is asked for), and freed ranges coalesce with their free neighbours.  Plain
// C++ (no device code), so host tests can drive it directly.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>

namespace rh {

class RangeAlloc {
   public:
    explicit RangeAlloc(uint64_t capacity = 0) { reset(capacity); }
    void reset(uint64_t capacity);
    bool alloc(uint64_t size, uint64_t* start, uint64_t align = 1);
    void free(uint64_t start, uint64_t size);
    uint64_t capacity() const { return cap_; }
    uint64_t used() const { return used_; }
    uint64_t high_water() const;  // end of the highest allocated range
    uint64_t largest_free() const;
    size_t free_ranges() const { return free_.size(); }

   private:
    std::map<uint64_t, uint64_t> free_;  // start -> size
    std::map<uint64_t, uint64_t> taken_;
    uint64_t cap_ = 0, used_ = 0;
};

}  // namespace rh
