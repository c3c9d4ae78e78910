// fc_table.h -- colour table of an enumerated false-colour subject.
//
// EnumeratedFCSubject (FCSubject.h:153-167) gives every id of a category of
// `size` ids its own colour: `size` distinct 24-bit colours in
// [1, 0xFFFFFF] drawn uniformly, kept in ascending order (a std::set), then
// shuffled; an id >= size paints black.  The reference draws from an
// unseeded generator and shuffles with std::random_shuffle, so its tables
// change from run to run.  Here the draws and the shuffle come from
// splitmix64 (Steele, Lea and Flood 2014) started at (seed << 8 | subject):
// v = 1 + (x mod 0xFFFFFF) until `size` distinct values, ascending sort,
// then Fisher-Yates from the last element down with j = x mod (i + 1).
// The same seed and subject always give the same table.
#ifndef OMV_FC_TABLE_H
#define OMV_FC_TABLE_H

#include <algorithm>
#include <cstdint>
#include <vector>

namespace omv {

inline uint64_t splitmix64Next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// false when size exceeds the 0xFFFFFF distinct colours there are (the
// reference's draw loop would not end)
inline bool fcColorTable(uint32_t seed, int subject, uint32_t size, std::vector<uint32_t>& out) {
    out.clear();
    if (size > 0xFFFFFFu) return false;
    uint64_t s = ((uint64_t)seed << 8) | (uint64_t)(subject & 0xFF);
    std::vector<uint8_t> seen((0xFFFFFFu >> 3) + 1, 0);
    out.reserve(size);
    while (out.size() < size) {
        const uint32_t v = 1u + (uint32_t)(splitmix64Next(s) % 0xFFFFFFull);
        if (seen[v >> 3] & (1u << (v & 7))) continue;
        seen[v >> 3] = (uint8_t)(seen[v >> 3] | (1u << (v & 7)));
        out.push_back(v);
    }
    std::sort(out.begin(), out.end());
    for (uint32_t i = size; i > 1; --i) {
        const uint32_t j = (uint32_t)(splitmix64Next(s) % i);
        std::swap(out[i - 1], out[j]);
    }
    return true;
}

}  // namespace omv

#endif
