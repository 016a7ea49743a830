// Host-only check of the segmented-sort plan tables (compute_amd/csrc/
// segplan.hpp), built by tests/test_sanitizers.py with AddressSanitizer and
// UndefinedBehaviorSanitizer.  For random and edge-case segment lists it
// checks that every segment of two or more elements lands in exactly one
// path with its own base and length, that class capacities hold, that the
// tiles and histogram chunks of every large segment cover it exactly once
// in order, and that bad offsets are refused.  Prints "<n> failures".
#include "../../compute_amd/csrc/segplan.hpp"

#include <cstdio>
#include <algorithm>
#include <map>
#include <random>
#include <string>
#include <vector>

using namespace bcmi;

static int failures = 0;
#define CHECK(c)                                                                    \
    do {                                                                            \
        if (!(c)) {                                                                 \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);                \
            ++failures;                                                             \
        }                                                                           \
    } while (0)

static void check_plan(const std::vector<uint64_t> &off, size_t eb, bool flt, uint64_t tile, int cls_cap = 4)
{
    const size_t nseg = off.size() - 1;
    SegTables t;
    std::string err;
    CHECK(build_seg_tables(off.data(), nseg, eb, flt, tile, t, err, cls_cap) == 0);
    CHECK(t.total == off[nseg]);
    CHECK(t.small_cls >= 0 && t.small_cls <= 4 && t.small_cls <= std::max(cls_cap, 0));
    const uint64_t small_max = (uint64_t)kSegSmallThreads << t.small_cls;
    // which path each segment took, by base
    std::map<uint64_t, int> seen;
    for (int c = 0; c < 5; ++c)
        for (const SegSpan &s : t.small[c]) {
            CHECK(c <= t.small_cls);
            CHECK(s.len <= ((uint64_t)kSegSmallThreads << c));
            CHECK(c == 0 || s.len > ((uint64_t)kSegSmallThreads << (c - 1)));
            seen[s.base] += 1;
        }
    for (const SegSpan &s : t.insert) {
        CHECK(flt && s.len >= 2 && s.len < 32);
        seen[s.base] += 1;
    }
    CHECK(t.chunk_begin.size() == t.large.size() + 1 && t.slot_begin.size() == t.large.size() + 1);
    size_t ti = 0;
    for (size_t li = 0; li < t.large.size(); ++li) {
        const SegSpan &s = t.large[li];
        CHECK(s.idx == ti && s.len > small_max);
        seen[s.base] += 1;
        // tiles: 0, 1, ... covering [0, len)
        const uint64_t ntiles = (s.len + tile - 1) / tile;
        for (uint64_t k = 0; k < ntiles; ++k, ++ti) {
            CHECK(ti < t.tiles.size());
            if (ti >= t.tiles.size()) return;
            const SegTile &x = t.tiles[ti];
            CHECK(x.base == s.base && x.len == s.len && x.tile == k && x.seg == li);
        }
        // chunks: consecutive pieces of kSegChunk covering the segment
        uint64_t pos = s.base;
        for (uint32_t c = t.chunk_begin[li]; c < t.chunk_begin[li + 1]; ++c) {
            CHECK(c < t.chunks.size());
            if (c >= t.chunks.size()) return;
            CHECK(t.chunks[c].base == pos && t.chunks[c].idx == li && t.chunks[c].len <= kSegChunk &&
                  t.chunks[c].len > 0);
            pos += t.chunks[c].len;
        }
        CHECK(pos == s.base + s.len);
        CHECK(t.slot_begin[li + 1] > t.slot_begin[li] && t.slot_begin[li + 1] - t.slot_begin[li] <= 256);
    }
    CHECK(ti == t.tiles.size());
    CHECK(t.chunk_begin.back() == t.chunks.size());
    // every segment of >= 2 elements exactly once, shorter ones never
    for (size_t s = 0; s < nseg; ++s) {
        const uint64_t len = off[s + 1] - off[s];
        auto it = seen.find(off[s]);
        if (len >= 2) {
            CHECK(it != seen.end() && it->second == 1);
        } else if (it != seen.end()) {
            // another segment starts at the same base only if this one is empty
            CHECK(len == 0);
        }
    }
    size_t listed = 0;
    for (auto &kv : seen) listed += kv.second;
    size_t expect = 0;
    for (size_t s = 0; s < nseg; ++s) expect += (off[s + 1] - off[s]) >= 2;
    CHECK(listed == expect);
}

int main()
{
    std::mt19937_64 g(2026);
    const uint64_t tiles[] = {4096, 3072, 2048};
    const size_t ebs[] = {4, 8, 12, 16, 24};
    // edge cases
    check_plan({0}, 12, false, 4096);                                // no segments
    check_plan({0, 0, 1, 1, 3}, 12, true, 4096);                    // empty and single-element segments
    check_plan({0, 4096, 4097 + 4096, 4097 + 4096 + (1ull << 20)}, 4, false, 4096);
    check_plan({0, 31, 63, 95}, 8, true, 3072);                     // float insertion-sort segments
    check_plan({0, (1ull << 26) + 5}, 12, false, 3072);              // one large segment, ragged tail
    for (int rep = 0; rep < 300; ++rep) {
        const size_t nseg = 1 + g() % 3000;
        std::vector<uint64_t> off(nseg + 1, 0);
        for (size_t s = 0; s < nseg; ++s) {
            uint64_t len;
            switch (g() % 6) {
            case 0: len = g() % 3; break;
            case 1: len = g() % 40; break;
            case 2: len = g() % 5000; break;
            case 3: len = 4096 + g() % 70000; break;
            case 4: len = (g() % 8 == 0) ? (1ull << 20) + g() % 1000 : g() % 300; break;
            default: len = g() % 1200;
            }
            off[s + 1] = off[s] + len;
        }
        check_plan(off, ebs[rep % 5], rep % 2 == 0, tiles[rep % 3], 4 - rep % 5);
    }
    // refused: decreasing offsets, an over-long segment
    {
        std::vector<uint64_t> off = {0, 10, 5};
        SegTables t;
        std::string err;
        CHECK(build_seg_tables(off.data(), 2, 8, false, 4096, t, err) == BCMI_INVALID_VALUE && !err.empty());
        std::vector<uint64_t> big = {0, 1ull << 33};
        CHECK(build_seg_tables(big.data(), 1, 8, false, 4096, t, err) == BCMI_INVALID_BUFFER_SIZE);
    }
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
