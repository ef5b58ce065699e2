// dtb_ranges.hpp -- best-fit allocation of byte ranges within one segment,
// with neighbouring free ranges merged on release.  Host bookkeeping only.
#pragma once

#include <cstddef>
#include <iterator>
#include <map>

namespace dtb {

class RangeSet {
public:
    explicit RangeSet(size_t size = 0) : size_(size), free_(size)
    {
        if (size > 0)
            insert(0, size);
    }
    size_t size() const { return size_; }
    size_t free_bytes() const { return free_; }
    bool all_free() const { return free_ == size_; }
    // Smallest free range of at least len bytes, or 0 if none.
    size_t best(size_t len) const
    {
        auto it = by_len_.lower_bound(len);
        return it == by_len_.end() ? 0 : it->first;
    }
    // Takes len bytes from the front of the smallest free range that holds
    // them; false if none does.
    bool take(size_t len, size_t *off)
    {
        auto it = by_len_.lower_bound(len);
        if (len == 0 || it == by_len_.end())
            return false;
        const size_t have = it->first, at = it->second;
        erase(at, have);
        if (have > len)
            insert(at + len, have - len);
        free_ -= len;
        *off = at;
        return true;
    }
    // Returns [off, off + len), merging it with free neighbours.
    void give(size_t off, size_t len)
    {
        free_ += len;
        auto next = by_off_.lower_bound(off);
        if (next != by_off_.end() && off + len == next->first) {
            len += next->second;
            erase(next->first, next->second);
        }
        auto prev = by_off_.lower_bound(off);
        if (prev != by_off_.begin()) {
            --prev;
            if (prev->first + prev->second == off) {
                off = prev->first;
                len += prev->second;
                erase(prev->first, prev->second);
            }
        }
        insert(off, len);
    }

private:
    void insert(size_t off, size_t len)
    {
        by_off_[off] = len;
        by_len_.emplace(len, off);
    }
    void erase(size_t off, size_t len)
    {
        by_off_.erase(off);
        auto r = by_len_.equal_range(len);
        for (auto it = r.first; it != r.second; ++it)
            if (it->second == off) {
                by_len_.erase(it);
                break;
            }
    }
    size_t size_, free_;
    std::map<size_t, size_t> by_off_;        // offset -> length
    std::multimap<size_t, size_t> by_len_;   // length -> offset
};

}  // namespace dtb
