// bmi_arena.h -- host-side bookkeeping of the request server's HBM arena
// and its admission rule.  Pure C++ (no HIP): the server owns one device
// allocation and hands out offsets from it; the same code is unit-tested
// on the CPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <iterator>
#include <map>
#include <vector>

namespace bmi {

// First-fit allocator over [0, capacity) with coalescing of free ranges.
class Arena {
public:
    static constexpr size_t kNone = ~(size_t)0;

    explicit Arena(size_t capacity = 0) { reset(capacity); }

    void reset(size_t capacity)
    {
        free_.clear();
        used_.clear();
        capacity_ = capacity;
        free_bytes_ = capacity;
        if (capacity) free_[0] = capacity;
    }

    // offset of a block of `bytes` aligned to `align` (a power of two), or
    // kNone when no free range holds it
    size_t alloc(size_t bytes, size_t align = 256)
    {
        if (bytes == 0) bytes = 1;
        bytes = (bytes + align - 1) & ~(align - 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            const size_t start = (it->first + align - 1) & ~(align - 1);
            const size_t end = it->first + it->second;
            if (start + bytes > end) continue;
            const size_t fo = it->first;
            free_.erase(it);
            if (start > fo) free_[fo] = start - fo;
            if (start + bytes < end) free_[start + bytes] = end - (start + bytes);
            used_[start] = bytes;
            free_bytes_ -= bytes;
            return start;
        }
        return kNone;
    }

    // returns false for an offset that is not an allocated block
    bool release(size_t off)
    {
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t start = off, len = u->second;
        used_.erase(u);
        free_bytes_ += len;
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
        return true;
    }

    size_t capacity() const { return capacity_; }
    size_t free_bytes() const { return free_bytes_; }
    size_t used_blocks() const { return used_.size(); }
    size_t largest_free() const
    {
        size_t m = 0;
        for (const auto &f : free_) m = f.second > m ? f.second : m;
        return m;
    }

private:
    std::map<size_t, size_t> free_;  // offset -> length
    std::map<size_t, size_t> used_;  // offset -> length
    size_t capacity_ = 0;
    size_t free_bytes_ = 0;
};

// Admission rule of the server: requests are admitted strictly in arrival
// order (no overtaking, so a large request cannot starve), as many as fit
// together in `avail` bytes, at most `max_batch`.  Returns how many of the
// pending requests (their footprints in `bytes`, oldest first) to admit.
inline int admit_prefix(const std::vector<size_t> &bytes, size_t avail, int max_batch)
{
    size_t sum = 0;
    int n = 0;
    for (size_t b : bytes) {
        if (n >= max_batch || sum + b > avail) break;
        sum += b;
        ++n;
    }
    return n;
}

}  // namespace bmi
