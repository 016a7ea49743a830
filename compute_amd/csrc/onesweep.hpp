// onesweep.hpp -- one 8-bit digit pass of the LSD radix sort, ranked by
// quarter-waves.
//
// A workgroup of THREADS threads owns tile `blockIdx.x` of THREADS * IPT
// keys.  Each 16-lane quarter-wave ("sub") owns a contiguous run of 16 * IPT
// positions of the tile, loaded quarter-striped: slot j of lane l holds
// position sub * 16 * IPT + j * 16 + (l & 15), so every load instruction
// reads four 64-byte segments and ranking slot by slot, lane by lane, walks
// the sub's keys in input order (stability).
//
// Per sub and digit, one 32-bit LDS word [position:16 | peer mask:16]:
//   1. early counts: every key adds 1 << 16 to its (sub, digit) word;
//   2. a scan in digit-major, sub-minor order turns the counts into the tile
//      positions where each (sub, digit) run starts, and gives the tile's
//      per-digit totals, which are published for the look-back right away;
//   3. ranking, per slot: each lane ORs its lane bit into the low half of its
//      word, reads the word back (peers = same-digit lanes of the slot),
//      takes position = start + peers below it, and every peer writes the
//      same advanced start with the mask cleared.  LDS executes one wave's
//      instructions in order, so slot j + 1 sees slot j's update.
// The look-back fetches the global offset of each digit: its first reads are
// issued before ranking, and the tile's first staging round is written to
// LDS (reordered by position, digit runs contiguous) while it completes.
// Then the tile is written out coalesced.  Three LDS round trips per key to
// rank, one atomic add to count, one write and one read to reorder.
//
// LDS: the count words alias the staging buffer (ranking finishes before
// the first staging write), so a tile needs max(staged bytes, THREADS * 64)
// bytes plus 2-3 KB, the staged bytes being the whole tile's keys (or
// values) or, staged in SPLIT rounds, a round's keys and values.
#pragma once

#include "radix.hpp"

namespace bcmi {

// SPLIT = 1: the whole tile is staged, keys then values.  SPLIT > 1: the
// tile's positions are staged in SPLIT rounds of TILE / SPLIT, keys and
// values of a round side by side, which shrinks the LDS a workgroup holds
// (more workgroups per CU to hide memory latency) for one more barrier per
// round.
template <int THREADS, int IPT, typename KB, int VB, int SPLIT>
struct QSweepLds {
    static constexpr int NSUB = THREADS / 16;
    static constexpr int TILE = THREADS * IPT;
    static constexpr int H = TILE / SPLIT;
    static constexpr int SB = SPLIT == 1 ? ((int)sizeof(KB) > VB ? (int)sizeof(KB) : VB) * TILE
                                         : ((int)sizeof(KB) + VB) * H;
    static constexpr int WORDS = NSUB * 256 * 4;
    static constexpr int STAGE = SB > WORDS ? SB : WORDS;
    alignas(16) unsigned char stage[STAGE];
    uint32_t goff[256];
    uint32_t part[THREADS / 256][256];
    uint32_t wsum[4];
};

// x, with the compiler made to forget how it was computed (so values
// derived from it earlier are recomputed, not kept alive in registers).
template <typename KB>
BDEV KB opaque_reg(KB x)
{
    if constexpr (sizeof(KB) <= 4) {
        uint32_t u = (uint32_t)x;
        asm volatile("" : "+v"(u));
        return (KB)u;
    } else {
        uint64_t u = (uint64_t)x;
        asm volatile("" : "+v"(u));
        return (KB)u;
    }
}

// Tile positions (below 2^16), two per register.
template <int IPT>
struct Pos16 {
    uint32_t w[(IPT + 1) / 2];
    BDEV void set(int j, uint32_t v)
    {
        if (j & 1) {
            w[j >> 1] |= v << 16;
            // materialise the packed pair here (see Ranks in radix.hpp)
            asm volatile("" : "+v"(w[j >> 1]));
        } else {
            w[j >> 1] = v;
        }
    }
    BDEV uint32_t operator[](int j) const { return (j & 1) ? (w[j >> 1] >> 16) : (w[j >> 1] & 0xFFFFu); }
};

// Digits of up to 4 keys per register (the value pass reuses the key pass's
// digits).
template <int IPT>
struct Dig8 {
    uint32_t w[(IPT + 3) / 4];
    BDEV void set(int j, uint32_t d)
    {
        if ((j & 3) == 0) w[j >> 2] = d;
        else w[j >> 2] |= d << (8 * (j & 3));
    }
    BDEV uint32_t operator[](int j) const { return (w[j >> 2] >> (8 * (j & 3))) & 0xFFu; }
};

// Look-back for digit d of tile `tile` (tile > 0).  The first LB
// predecessor words are read early (before the tile ranks its keys, so the
// round trip overlaps the ranking) and handed in as wd; then the
// predecessors' aggregates are folded in tile order until an inclusive
// prefix, LB words per round trip; at a word not yet published only that
// word is polled.  A word read early may be older than the predecessor's
// latest state, which is harmless: an aggregate instead of an inclusive
// prefix only sends the walk further back.  Returns the exclusive prefix (a
// bounded wait: after kSpinLimit polls it records a fault and returns what
// it has).
template <int LB>
BDEV void digit_lookback_issue(const uint64_t *status, uint32_t tile, uint32_t d, uint32_t epoch,
                               uint64_t (&wd)[LB])
{
#pragma unroll
    for (int i = 0; i < LB; ++i) {
        const int64_t t = (int64_t)tile - 1 - i;
        wd[i] = t >= 0 ? lb_load(&status[(size_t)t * 256 + d]) : lb_word(epoch, LB_P, 0);
    }
}

template <int LB>
BDEV uint32_t digit_lookback(const uint64_t *status, uint32_t tile, uint32_t d, uint32_t epoch, Ctrl *ctrl,
                             uint64_t (&wd)[LB])
{
    uint32_t excl = 0, spins = 0;
    int64_t kt = (int64_t)tile - 1;
    while (true) {
        int used = 0;
        bool found = false, stalled = false;
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            if (found || stalled) continue;
            if (lb_epoch(wd[i]) != epoch) {
                stalled = true;
                continue;
            }
            excl += lb_payload(wd[i]);
            ++used;
            if (lb_flag(wd[i]) == LB_P) found = true;
        }
        if (found) return excl;
        kt -= used;
        if (stalled && used == 0) {
            while (lb_epoch(lb_load(&status[(size_t)kt * 256 + d])) != epoch)
                if (spin_expired(ctrl, spins)) return excl;
        }
        digit_lookback_issue<LB>(status, (uint32_t)(kt + 1), d, epoch, wd);
    }
}

template <int THREADS, int IPT, typename KB, int VB, int SPLIT, int LB, typename XF, bool FULL, int CBG>
BDEV void qsweep_body(QSweepLds<THREADS, IPT, KB, VB, SPLIT> &L, const KB *__restrict__ kin, KB *__restrict__ kout,
                      const typename ValT<VB>::type *__restrict__ vin, typename ValT<VB>::type *__restrict__ vout,
                      uint32_t n, uint32_t tile, uint32_t shift, const uint32_t *__restrict__ goff0, XF xf, Ctrl *ctrl,
                      uint64_t *__restrict__ status, uint32_t epoch)
{
    using V = typename ValT<VB>::type;
    constexpr int NSUB = THREADS / 16, TILE = THREADS * IPT, PARTS = THREADS / 256, SPP = NSUB / PARTS;
    static_assert(THREADS % 256 == 0 && TILE <= 65536, "256-multiple blocks, tile positions below 2^16");
    static_assert(IPT % SPLIT == 0, "whole slots per staging round");
    uint32_t *words = (uint32_t *)L.stage;
    KB *s_keys = (KB *)L.stage;
    V *s_vals = (V *)L.stage;
    const uint32_t tid = threadIdx.x, ql = tid & 15, sub = tid >> 4;
    // Count words, 256 per sub.  CBG = 0: sub s's digit d at s * 256 + d, so
    // digit d sits in bank d mod 64 in every sub and the 64 lanes of a wave
    // contend for the same 64 banks.  CBG = 1 (bank groups): the 4 subs of a
    // wave share 1024 words as 16 rows of 64, sub sw owning columns
    // [16 sw, 16 sw + 16), digit d at row d / 16, column 16 sw + d % 16, so a
    // sub's 16 lanes contend among 16 banks of their own.
    auto cw = [](uint32_t s, uint32_t d) -> uint32_t {
        if constexpr (CBG) return (s >> 2) * 1024 + (d >> 4) * 64 + (s & 3) * 16 + (d & 15);
        else return s * 256 + d;
    };
    // each sub zeroes its own 256 count words (16 per lane, four 16-byte
    // stores); only the sub's lanes touch them before the barrier that ends
    // the counting, and one wave's LDS operations execute in order
    uint32_t *my = words + cw(sub, 0);
    auto mw = [&](uint32_t d) -> uint32_t * {
        if constexpr (CBG) return my + (d >> 4) * 64 + (d & 15);
        else return my + d;
    };
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if constexpr (CBG) *(uint4 *)(my + (i * 4 + (ql >> 2)) * 64 + (ql & 3) * 4) = make_uint4(0, 0, 0, 0);
        else ((uint4 *)my)[i * 16 + ql] = make_uint4(0, 0, 0, 0);
    }
    const uint32_t base = tile * (uint32_t)TILE;
    const int32_t lim = FULL ? TILE : (int32_t)(n - base) - (int32_t)(sub * 16 * IPT);
    auto valid = [&](int j) { return FULL || (int32_t)(j * 16 + ql) < lim; };
    const KB *kq = kin + ((size_t)base + sub * 16 * IPT + ql);
    KB k[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = valid(j) ? kq[j * 16] : KB(0);
    V v[VB ? IPT : 1];
    if constexpr (VB > 0) {
        const V *vq = vin + ((size_t)base + sub * 16 * IPT + ql);
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (valid(j)) v[j] = vq[j * 16];
    }
    // 1. early counts
#pragma unroll
    for (int j = 0; j < IPT; ++j)
        if (valid(j))
            __hip_atomic_fetch_add(mw(xf.digit(k[j], shift)), 1u << 16, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
    block_barrier();
    // 2. thread (part, d) scans digit d over subs [part * SPP, part * SPP + SPP)
    const uint32_t d = tid & 255, part = tid >> 8;
    uint32_t c[SPP];
    uint32_t psum = 0;
#pragma unroll
    for (int i = 0; i < SPP; ++i) {
        c[i] = words[cw(part * SPP + i, d)] >> 16;
        psum += c[i];
    }
    uint32_t total = psum, before = 0;
    if constexpr (PARTS > 1) {
        L.part[part][d] = psum;
        block_barrier();
        total = 0;
#pragma unroll
        for (int p = 0; p < PARTS; ++p) {
            const uint32_t x = L.part[p][d];
            if (p < (int)part) before += x;
            total += x;
        }
    }
    const uint32_t lane = tid & 63, w = tid >> 6;
    const uint32_t inc = wave_inclusive_scan(total, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if (lane == 63 && w < 4) L.wsum[w] = inc;
    if (part == 0) {
        if (tile == 0) lb_store(&status[d], lb_word(epoch, LB_P, total));
        else lb_store(&status[(size_t)tile * 256 + d], lb_word(epoch, LB_A, total));
    }
    block_barrier();
    uint32_t dstart = inc - total;
    for (uint32_t i = 0; i < (w & 3); ++i) dstart += L.wsum[i];
    {
        uint32_t s0 = dstart + before;
#pragma unroll
        for (int i = 0; i < SPP; ++i) {
            words[cw(part * SPP + i, d)] = s0 << 16;
            s0 += c[i];
        }
    }
    block_barrier();
    // the look-back's first reads, in flight while the tile ranks
    uint64_t lbw[LB];
    if (part == 0 && tile > 0) digit_lookback_issue<LB>(status, tile, d, epoch, lbw);
    // 3. rank
    Pos16<IPT> pos;
    const uint32_t qbit = 1u << ql, qlt = qbit - 1u;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool ok = valid(j);
        // the digit is recomputed here rather than kept from the counting
        // loop: keeping all IPT word addresses alive across the scan would
        // cost IPT more VGPRs per lane (occupancy is register-bound)
        uint32_t *a = mw(xf.digit(opaque_reg(k[j]), shift));
        if (ok) __hip_atomic_fetch_or(a, qbit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t wd = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t run = wd >> 16;
        pos.set(j, run + __builtin_popcount(wd & qlt));
        if (ok)
            __hip_atomic_store(a, (run + __builtin_popcount(wd & 0xFFFFu)) << 16, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    // 4. reorder by position (digit runs contiguous) and write out
    // coalesced, in SPLIT rounds of H positions (keys and values of a round
    // side by side), or, with SPLIT = 1, keys then values.  The first round
    // is staged while the look-back completes; the count words share the
    // staging buffer, so every sub must have finished ranking first.
    block_barrier();
    constexpr int H = TILE / SPLIT;
    V *s_v = (V *)(L.stage + (SPLIT > 1 ? H * sizeof(KB) : 0));
    auto stage_round = [&](int r) {
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t p = pos[j] - (uint32_t)(r * H);
            if (valid(j) && p < (uint32_t)H) {
                s_keys[p] = k[j];
                if constexpr (VB > 0 && SPLIT > 1) s_v[p] = v[j];
            }
        }
    };
    stage_round(0);
    // global offsets: digit d's keys go to [goff0 + excl, goff0 + excl + total),
    // checked once per digit; a tile with an inconsistent offset (a corrupted
    // chain) writes nothing and reports
    bool bad = false;
    if (part == 0) {
        uint32_t excl = 0;
        if (tile > 0) {
            excl = digit_lookback<LB>(status, tile, d, epoch, ctrl, lbw);
            lb_store(&status[(size_t)tile * 256 + d], lb_word(epoch, LB_P, excl + total));
        }
        L.goff[d] = goff0[d] + excl - dstart;
        bad = (uint64_t)goff0[d] + excl + total > n;
    }
    if (block_barrier_or(bad)) {
        if (tid == 0) __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const uint32_t tile_n = FULL ? (uint32_t)TILE : n - base;
    if constexpr (SPLIT > 1) {
        constexpr int JR = IPT / SPLIT;
#pragma unroll
        for (int r = 0; r < SPLIT; ++r) {
            if (r > 0) {
                block_barrier();  // the previous round's reads are done
                stage_round(r);
                block_barrier();
            }
#pragma unroll
            for (int jj = 0; jj < JR; ++jj) {
                const uint32_t li = jj * THREADS + tid, i = (uint32_t)(r * H) + li;
                if (FULL || i < tile_n) {
                    const KB key = s_keys[li];
                    const uint32_t o = L.goff[xf.digit(key, shift)] + i;
                    kout[o] = key;
                    if constexpr (VB > 0) vout[o] = s_v[li];
                }
            }
        }
        return;
    }
    Dig8<(VB > 0 ? IPT : 1)> od;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const uint32_t i = j * THREADS + tid;
        if (FULL || i < tile_n) {
            const KB key = s_keys[i];
            const uint32_t dg = xf.digit(key, shift);
            if constexpr (VB > 0) od.set(j, dg);
            kout[L.goff[dg] + i] = key;
        }
    }
    if constexpr (VB > 0) {
        block_barrier();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (valid(j)) s_vals[pos[j]] = v[j];
        block_barrier();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t i = j * THREADS + tid;
            if (FULL || i < tile_n) vout[L.goff[od[j]] + i] = s_vals[i];
        }
    }
}

// Tile of workgroup b.  Workgroups go to the XCDs round-robin (b mod 8), so
// with xrun = 0 consecutive tiles run on different XCDs and the partial
// 64-byte lines where one tile's digit run ends and the next tile's begins
// reach HBM from two L2s.  xrun = R > 0 renumbers each group of 8R
// workgroups so that tiles [g*8R + x*R, g*8R + x*R + R) run on XCD x and
// those lines meet in one L2; a trailing partial group keeps b.  A tile's
// look-back may then wait on a workgroup dispatched up to 8R - 1 later,
// which needs 8R workgroups resident at once (the host enables it only
// when the device holds far more).
BDEV uint32_t xcd_tile(uint32_t b, uint32_t xrun)
{
    if (xrun == 0) return b;
    const uint32_t G = 8 * xrun, g = b / G, r = b - g * G;
    if ((g + 1) * G > gridDim.x) return b;
    return g * G + (r & 7) * xrun + (r >> 3);
}

// WPE: the waves per SIMD the register allocation must leave room for (1 =
// no bound beyond the block size).  Occupancy of this kernel is bound by
// its VGPRs: a bound of 4 fits four 256-thread workgroups per CU in 128
// VGPRs, 8 fits four 512-thread workgroups in 64.
template <int THREADS, int IPT, typename KB, int VB, int SPLIT, int LB, int WPE, typename XF, int CBG = 0>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_qsweep(const KB *__restrict__ kin, KB *__restrict__ kout,
                                                    const typename ValT<VB>::type *__restrict__ vin,
                                                    typename ValT<VB>::type *__restrict__ vout, uint32_t n,
                                                    uint32_t shift, uint32_t pass, XF xf, Ctrl *ctrl,
                                                    uint64_t *__restrict__ status, uint32_t back, uint32_t xrun)
{
    __shared__ QSweepLds<THREADS, IPT, KB, VB, SPLIT> lds;
    const uint32_t tile = xcd_tile(blockIdx.x, xrun);
    const uint32_t epoch = launch_epoch_plain(ctrl, back);
    if ((uint64_t)(tile + 1) * (THREADS * IPT) <= n)
        qsweep_body<THREADS, IPT, KB, VB, SPLIT, LB, XF, true, CBG>(lds, kin, kout, vin, vout, n, tile, shift,
                                                           ctrl->offsets[pass], xf, ctrl, status, epoch);
    else
        qsweep_body<THREADS, IPT, KB, VB, SPLIT, LB, XF, false, CBG>(lds, kin, kout, vin, vout, n, tile, shift,
                                                            ctrl->offsets[pass], xf, ctrl, status, epoch);
}

}  // namespace bcmi
