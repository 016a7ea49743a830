// segplan.hpp -- host-only tables of a segmented-sort plan.
//
// From the host offsets of nseg segments, the plan routes every segment of
// two or more elements to exactly one path and lists its work items:
//   * small[c]: segments of at most kSegSmallThreads << c elements (c up to
//     the largest class whose key + value bytes fit the LDS budget), sorted
//     entirely in LDS by one workgroup each;
//   * insert:   float keys shorter than 32 (the reference's insertion sort);
//   * large:    the rest, each with its onesweep tiles (tile elements apiece)
//     and histogram chunks of kSegChunk elements, and a budget of group
//     slots for the MSD route.
// No HIP here: segmented.hip uploads the tables, and the CPU test suite
// builds them under AddressSanitizer / UndefinedBehaviorSanitizer.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bcmi.h"

namespace bcmi {

constexpr int kSegSmallThreads = 256;
constexpr int kSegSmallLds = 48 * 1024;   // key + value bytes a small-segment block keeps in LDS
constexpr uint32_t kSegChunk = 1u << 16;  // keys per histogram chunk of a large segment

struct SegSpan {
    uint64_t base;
    uint32_t len;
    uint32_t idx;  // large segments: index of the first tile; chunks: large-segment index; else unused
};

struct SegTile {
    uint64_t base;
    uint32_t len;
    uint32_t tile;   // tile index within the segment
    uint32_t seg;    // large-segment index
    uint32_t route;  // the segment's route, written on the device before the passes run
};

struct SegTables {
    int small_cls = 4;                 // largest small class
    uint64_t total = 0;                // elements covered by the offsets
    std::vector<SegSpan> small[5];     // classes of capacity 256, 512, ..., 4096
    std::vector<SegSpan> insert, large, chunks;
    std::vector<uint32_t> chunk_begin{0};  // per large segment: first chunk (nlarge + 1 entries)
    std::vector<uint32_t> slot_begin{0};   // per large segment: first MSD group slot (nlarge + 1)
    std::vector<SegTile> tiles;
};

// Largest small class for key_bytes + value_bytes per element.
inline int seg_small_class(size_t elem_bytes)
{
    int cmax = 4;
    while (cmax > 0 && elem_bytes * (size_t)kSegSmallThreads * ((size_t)1 << cmax) > (size_t)kSegSmallLds) --cmax;
    return cmax;
}

// Builds the tables; returns 0, or BCMI_INVALID_VALUE (offsets decrease) /
// BCMI_INVALID_BUFFER_SIZE (a segment too long) with a message in err.
// tile: elements per onesweep tile of a large segment; small_cls_cap: the
// largest small class allowed (the LDS budget may lower it further).
inline int build_seg_tables(const uint64_t *offsets, size_t nseg, size_t elem_bytes, bool float_keys,
                             uint64_t tile, SegTables &t, std::string &err, int small_cls_cap = 4)
{
    t = SegTables();
    t.small_cls = std::max(0, std::min(seg_small_class(elem_bytes), small_cls_cap));
    const uint64_t small_max = (uint64_t)kSegSmallThreads << t.small_cls;
    t.total = nseg ? offsets[nseg] : 0;
    char buf[160];
    for (size_t s = 0; s < nseg; ++s) {
        if (offsets[s + 1] < offsets[s]) {
            std::snprintf(buf, sizeof(buf), "segmented plan: offsets decrease at segment %zu", s);
            err = buf;
            return BCMI_INVALID_VALUE;
        }
        const uint64_t base = offsets[s], len = offsets[s + 1] - offsets[s];
        if (len > (1ull << 32) - (1ull << 20)) {
            std::snprintf(buf, sizeof(buf), "segmented plan: segment %zu has %llu elements", s,
                          (unsigned long long)len);
            err = buf;
            return BCMI_INVALID_BUFFER_SIZE;
        }
        if (len < 2) continue;
        if (float_keys && len < 32) {
            t.insert.push_back(SegSpan{base, (uint32_t)len, 0});
        } else if (len <= small_max) {
            int c = 0;
            while (((uint64_t)kSegSmallThreads << c) < len) ++c;
            t.small[c].push_back(SegSpan{base, (uint32_t)len, 0});
        } else {
            const uint32_t li = (uint32_t)t.large.size();
            t.large.push_back(SegSpan{base, (uint32_t)len, (uint32_t)t.tiles.size()});
            for (uint64_t k = 0; k * tile < len; ++k) t.tiles.push_back(SegTile{base, (uint32_t)len, (uint32_t)k, li, 0});
            for (uint64_t c = 0; c < len; c += kSegChunk) {
                const uint64_t cl = len - c < kSegChunk ? len - c : kSegChunk;
                t.chunks.push_back(SegSpan{base + c, (uint32_t)cl, li});
            }
            t.chunk_begin.push_back((uint32_t)t.chunks.size());
            // greedy groups of <= small_max keys: any two consecutive ones
            // hold more than small_max, so at most 2*len/small_max + 1
            const uint64_t gmax = std::min<uint64_t>(256, 2 * (len / small_max) + 2);
            t.slot_begin.push_back(t.slot_begin.back() + (uint32_t)gmax);
        }
    }
    return 0;
}

}  // namespace bcmi
