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

bool RangeAlloc::alloc(uint64_t size, uint64_t* start) {
    if (size == 0) size = 1;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= size) {
            uint64_t s = it->first, rest = it->second - size;
            free_.erase(it);
            if (rest) free_[s + size] = rest;
            taken_[s] = size;
            used_ += size;
            *start = s;
            return true;
        }
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
