// The batched engine's first-fit range allocator (range_alloc.h).
#include "range_alloc.h"

#include <iterator>

namespace rh {

void RangeAlloc::reset(uint64_t capacity) {
    free_.clear();
    taken_.clear();
    cap_ = capacity;
    used_ = 0;
    if (capacity) free_[0] = capacity;
}

bool RangeAlloc::alloc(uint64_t size, uint64_t* start, uint64_t align) {
    if (size == 0) size = 1;
    if (align == 0) align = 1;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        const uint64_t s = it->first, len = it->second;
        const uint64_t a = (s + align - 1) / align * align;  // the range's first aligned unit
        if (a - s > len || len - (a - s) < size) continue;
        const uint64_t rest = len - (a - s) - size;
        free_.erase(it);
        if (a > s) free_[s] = a - s;
        if (rest) free_[a + size] = rest;
        taken_[a] = size;
        used_ += size;
        *start = a;
        return true;
    }
    return false;
}

void RangeAlloc::free(uint64_t start, uint64_t size) {
    auto t = taken_.find(start);
    if (t == taken_.end()) return;
    size = t->second;  // the range as it was handed out
    taken_.erase(t);
    used_ -= size;
    auto it = free_.emplace(start, size).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free_.erase(next);
    }
    if (it != free_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free_.erase(it);
        }
    }
}

uint64_t RangeAlloc::high_water() const {
    if (taken_.empty()) return 0;
    auto last = std::prev(taken_.end());
    return last->first + last->second;
}

uint64_t RangeAlloc::largest_free() const {
    uint64_t m = 0;
    for (const auto& kv : free_) m = kv.second > m ? kv.second : m;
    return m;
}

}  // namespace rh
