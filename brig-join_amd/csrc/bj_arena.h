// bj_arena.h -- first-fit allocator over the serving arena's byte range.
//
// The serving loop (bj_serve.cpp) carves every admitted request's workspace
// and rows out of one device arena with this allocator; it only does offset
// arithmetic on the host, so it is also exercised by the host-only checks.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <iterator>
#include <map>

namespace bj {

// First-fit allocator over [0, size), in multiples of ALIGN bytes, coalescing
// neighbouring free ranges on release.  Not thread-safe (the server calls it
// under its lock).
class ArenaAlloc {
public:
    static constexpr uint64_t ALIGN = 256;
    static uint64_t round(uint64_t bytes) { return (std::max<uint64_t>(bytes, 1) + ALIGN - 1) / ALIGN * ALIGN; }

    void reset(uint64_t size)
    {
        free_.clear();
        size = size / ALIGN * ALIGN;
        if (size) free_[0] = size;
        size_ = size;
        used_ = peak_ = 0;
    }
    bool alloc(uint64_t bytes, uint64_t& off)
    {
        bytes = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const uint64_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                peak_ = std::max(peak_, used_);
                return true;
            }
        }
        return false;
    }
    // Releases a range returned by alloc(bytes) (the same byte count).
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = round(bytes);
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
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }
    size_t free_ranges() const { return free_.size(); }
    // largest free range
    uint64_t largest() const
    {
        uint64_t m = 0;
        for (const auto& f : free_) m = std::max(m, f.second);
        return m;
    }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0, used_ = 0, peak_ = 0;
};

}  // namespace bj
