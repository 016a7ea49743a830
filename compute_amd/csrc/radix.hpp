// radix.hpp -- pieces shared by the radix sort and the segmented sort:
// the radix() key transform, value carriers, and in-wave digit ranking.
#pragma once

#include "device.hpp"

#include <type_traits>

namespace bcmi {

constexpr int kSortThreads = 256;
constexpr int kHistThreads = 256;
constexpr int kLookback = 8;  // predecessor status words fetched per look-back step

// radix() key transform (radix_sort.hpp:98-129) on the W-bit unsigned image
// x of the key, in one of four forms fixed at compile time (a kernel is
// instantiated per form, so the per-key cost is one or two ALU ops):
//   kXfXor     key = x ^ cx        unsigned ascending (cx = 0), unsigned
//                                  descending (cx = ~0: (T)(-1)-x == ~x),
//                                  signed ascending (cx = signbit)
//   kXfNegXor  key = (-x) ^ cx     signed descending (cx = signbit)
//   kXfFlt     key = x ^ (-(x >> msb) | signbit)        float ascending
//   kXfNegFlt  key = (-x) ^ (-(x >> msb) | signbit)     float descending
// (the last one keeps the reference's collisions of +/-0 with -/+denorm_min)
enum { kXfXor = 0, kXfNegXor = 1, kXfFlt = 2, kXfNegFlt = 3 };

template <typename KB, int XM>
struct KeyXf {
    KB cx;
    BHD KB key(KB x) const
    {
        constexpr int W = sizeof(KB) * 8;
        using SK = std::make_signed_t<KB>;
        const KB t = (XM == kXfNegXor || XM == kXfNegFlt) ? (KB)(KB(0) - x) : x;
        if constexpr (XM == kXfFlt || XM == kXfNegFlt)
            return (KB)(t ^ (KB)((KB)((SK)x >> (W - 1)) | (KB)(KB(1) << (W - 1))));
        else
            return (KB)(t ^ cx);
    }
    BHD uint32_t digit(KB x, uint32_t shift) const { return (uint32_t)(key(x) >> shift) & 0xFFu; }
};

// Host: call f(KeyXf<KB, form>{cx}) for the form that serves a key of kind
// (0 unsigned, 1 signed, 2 floating) in the given direction.  Floating forms
// exist only for 4- and 8-byte keys.
template <typename KB, typename F>
inline int with_key_xf(int kind, int descending, F &&f)
{
    const KB sgn = (KB)((KB)1 << (sizeof(KB) * 8 - 1));
    if constexpr (sizeof(KB) >= 4) {
        if (kind == 2) return descending ? f(KeyXf<KB, kXfNegFlt>{0}) : f(KeyXf<KB, kXfFlt>{0});
    }
    if (kind == 1) return descending ? f(KeyXf<KB, kXfNegXor>{sgn}) : f(KeyXf<KB, kXfXor>{sgn});
    return f(KeyXf<KB, kXfXor>{descending ? (KB)~(KB)0 : (KB)0});
}

template <int VB>
struct ValT;
template <> struct ValT<0> { using type = uint8_t; };
template <> struct ValT<1> { using type = uint8_t; };
template <> struct ValT<2> { using type = uint16_t; };
template <> struct ValT<4> { using type = uint32_t; };
template <> struct ValT<8> { using type = uint64_t; };
// 16-byte payloads as a native 4 x u32 vector: it lives in registers (a
// struct of two u64 was kept in scratch memory by the compiler)
typedef uint32_t v16_t __attribute__((ext_vector_type(4)));
template <> struct ValT<16> { using type = v16_t; };

// ------------------------------------------------------------ ranking
// Rank the IPT keys a lane holds (wave-striped: slot j of lane l is position
// j*64 + l of the wave's part of the tile) within the wave: r[j] = number of
// earlier keys of the wave with the same digit.  cnt/peer are the wave's 256
// digit counters (zeroed by the caller) and 256 peer-mask slots (zeroed by
// the caller, left zeroed).  Slot j of this lane is valid iff FULL or
// j*64 + lane < lim.
//   kRankLdsMask: each lane ORs its lane bit into its digit's slot
//     (ds_or_b64) and reads the slot back: the peer mask;
//   kRankBallot: 8 ballots, one per digit bit, AND-ed per lane.
// rank = wave counter + peers below this lane.  Every peer then writes the
// same new counter value (old + peers) and clears the slot, so no lane has
// to be singled out; LDS executes one wave's instructions in order, so slot
// j+1 sees slot j's updates.  Slots are ranked in order and lanes in lane
// order, i.e. in input order: the ranking is stable.
// In-wave ranks of a lane's IPT slots, two 16-bit ranks per register (a rank
// is below the tile size, at most 8192), which keeps the 32-keys-per-thread
// tile within 128 VGPRs.
template <int IPT>
struct Ranks {
    uint32_t w[(IPT + 1) / 2];
    BDEV void set(int j, uint32_t v)
    {
        if (j & 1) {
            w[j >> 1] |= v << 16;
            // materialise the packed pair here: left alone, the compiler
            // sinks the packing to the use after the barriers and keeps
            // both halves' operands live, which costs more registers
            asm volatile("" : "+v"(w[j >> 1]));
        } else {
            w[j >> 1] = v;
        }
    }
    BDEV uint32_t operator[](int j) const { return (j & 1) ? (w[j >> 1] >> 16) : (w[j >> 1] & 0xFFFFu); }
};

// dig(key) gives the digit in [0, 256).
template <int IPT, bool FULL, int RANK, typename KB, typename DigitF>
BDEV void rank_in_wave_by(const KB (&k)[IPT], Ranks<IPT> &r, DigitF dig, uint32_t *cnt, uint64_t *peer, int32_t lim)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lanebit = 1ull << lane;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool ok = FULL || (int32_t)(j * 64 + lane) < lim;
        const uint32_t d = dig(k[j]);
        uint64_t peers;
        if constexpr (RANK == 0) {
            if (ok) __hip_atomic_fetch_or(&peer[d], lanebit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            peers = __hip_atomic_load(&peer[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            peers = FULL ? ~0ull : ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint32_t nb = ((d >> b) & 1u) - 1u;  // 0 if the bit is set, ~0 if clear
                const uint64_t m = ballot(nb == 0u);
                peers &= m ^ (((uint64_t)nb << 32) | nb);
            }
        }
        const uint32_t old = cnt[d];
        r.set(j, old + mbcnt(peers));
        if (ok) {
            cnt[d] = old + popc64(peers);
            if constexpr (RANK == 0)
                __hip_atomic_store(&peer[d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

template <int IPT, bool FULL, int RANK, typename KB, typename XF>
BDEV void rank_in_wave(const KB (&k)[IPT], Ranks<IPT> &r, const XF &xf, uint32_t shift, uint32_t *cnt,
                       uint64_t *peer, int32_t lim)
{
    rank_in_wave_by<IPT, FULL, RANK>(k, r, [&](KB x) { return xf.digit(x, shift); }, cnt, peer, lim);
}

// Block-wide exclusive scan of one value per thread (256 threads).
BDEV uint32_t block_exclusive_sum256(uint32_t v, uint32_t *s_wsum)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if (lane == 63) s_wsum[w] = inc;
    block_barrier();
    uint32_t pre = inc - v;
    for (uint32_t i = 0; i < w; ++i) pre += s_wsum[i];
    return pre;
}

// ------------------------------------------------------------ onesweep pass
// RANK selects rank_in_wave's method (kRankLdsMask: ~12 VALU + 5 LDS ops
// per key; kRankBallot: ~40 VALU/SALU ops, no LDS round trip).
enum { kRankLdsMask = 0, kRankBallot = 1 };

// LDS of one onesweep tile.  The peer-mask slots of the ranking live in the
// staging buffer, which is only needed after ranking: 38 KB at 32 4-byte keys
// per thread, so four workgroups fit a CU.
template <typename KB, int VB, int IPT>
struct OnesweepLds {
    static constexpr int TILE = kSortThreads * IPT;
    static constexpr int SB = (int)sizeof(KB) > VB ? (int)sizeof(KB) : VB;
    static constexpr int STAGE = TILE * SB > 4 * 256 * 8 ? TILE * SB : 4 * 256 * 8;
    uint32_t cnt[4][256];
    uint32_t goff[256];
    uint32_t wsum[4];
    alignas(16) unsigned char stage[STAGE];
};

// One onesweep tile: keys [tile*TILE, tile*TILE + TILE) of an n-key chain.
// status points at the chain's first tile; goff0 holds the chain's 256
// exclusive digit offsets for this pass.  FULL: the tile lies entirely below
// n (every tile but a chain's last), so no per-key bounds test is needed.
template <typename KB, int VB, int IPT, int RANK, typename XF, bool FULL>
BDEV void onesweep_body(OnesweepLds<KB, VB, IPT> &L, const KB *__restrict__ kin, KB *__restrict__ kout,
                        const typename ValT<VB>::type *__restrict__ vin, typename ValT<VB>::type *__restrict__ vout,
                        uint32_t n, uint32_t tile, uint32_t shift, const uint32_t *__restrict__ goff0, XF xf,
                        Ctrl *ctrl, uint64_t *__restrict__ status, uint32_t back)
{
    using V = typename ValT<VB>::type;
    constexpr int TILE = kSortThreads * IPT;
    uint64_t(*s_peer)[256] = (uint64_t(*)[256])L.stage;
    KB *s_keys = (KB *)L.stage;
    V *s_vals = (V *)L.stage;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int i = tid; i < 4 * 256; i += kSortThreads) {
        (&L.cnt[0][0])[i] = 0;
        (&s_peer[0][0])[i] = 0;
    }
    block_barrier();
    const uint32_t base = tile * (uint32_t)TILE;
    // slot j of this lane (tile position w*64*IPT + j*64 + lane) is valid iff j*64 + lane < lim
    const int32_t lim = FULL ? TILE : (int32_t)(n - base) - (int32_t)(w * 64 * IPT);
    auto valid = [&](int j) { return FULL || (int32_t)(j * 64 + lane) < lim; };

    // wave-striped load: lane l of wave w holds tile positions w*64*IPT + j*64 + l
    // (through a wave-uniform base: one 32-bit lane offset + immediates)
    const uint32_t wu = __builtin_amdgcn_readfirstlane(w);
    const KB *kw = kin + ((size_t)base + wu * 64 * IPT);
    const V *vw = VB ? vin + ((size_t)base + wu * 64 * IPT) : vin;
    KB k[IPT];
    V v[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (FULL) {
            k[j] = kw[j * 64 + lane];
            if constexpr (VB > 0) v[j] = vw[j * 64 + lane];
        } else {
            k[j] = valid(j) ? kw[j * 64 + lane] : KB(0);
            if constexpr (VB > 0) {
                if (valid(j)) v[j] = vw[j * 64 + lane];
            }
        }
    }
    Ranks<IPT> r;
    rank_in_wave<IPT, FULL, RANK>(k, r, xf, shift, L.cnt[w], s_peer[w], lim);
    block_barrier();

    // per digit (thread tid == digit): wave prefixes and the tile count
    const uint32_t c0 = L.cnt[0][tid], c1 = L.cnt[1][tid], c2 = L.cnt[2][tid], c3 = L.cnt[3][tid];
    const uint32_t total = c0 + c1 + c2 + c3;
    const uint32_t epoch = launch_epoch_plain(ctrl, back);
    if (tile == 0) lb_store(&status[tid], lb_word(epoch, LB_P, total));
    else lb_store(&status[(size_t)tile * 256 + tid], lb_word(epoch, LB_A, total));
    uint32_t inc = wave_inclusive_scan(total, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if (lane == 63) L.wsum[w] = inc;
    block_barrier();
    uint32_t lstart = inc - total;
    for (uint32_t i = 0; i < w; ++i) lstart += L.wsum[i];
    L.cnt[0][tid] = lstart;
    L.cnt[1][tid] = lstart + c0;
    L.cnt[2][tid] = lstart + c0 + c1;
    L.cnt[3][tid] = lstart + c0 + c1 + c2;
    // decoupled look-back for this digit: fetch kLookback predecessors'
    // status words at once (one round trip), fold them in tile order until
    // an inclusive prefix; at a word not yet published, poll that word alone
    // until it is, then continue.
    uint32_t excl = 0;
    if (tile > 0) {
        int64_t kt = (int64_t)tile - 1;
        uint32_t spins = 0;
        while (true) {
            uint64_t wd[kLookback];
#pragma unroll
            for (int i = 0; i < kLookback; ++i)
                wd[i] = kt - i >= 0 ? lb_load(&status[(size_t)(kt - i) * 256 + tid]) : lb_word(epoch, LB_P, 0);
            int used = 0;
            bool found = false, stalled = false;
#pragma unroll
            for (int i = 0; i < kLookback; ++i) {
                if (found || stalled) continue;
                if (lb_epoch(wd[i]) != epoch) {
                    stalled = true;
                    continue;
                }
                excl += lb_payload(wd[i]);
                ++used;
                if (lb_flag(wd[i]) == LB_P) found = true;
            }
            if (found) break;
            kt -= used;
            if (stalled && used == 0) {
                // poll just the unpublished word, then read 8-wide again
                bool expired = false;
                while (lb_epoch(lb_load(&status[(size_t)kt * 256 + tid])) != epoch)
                    if ((expired = spin_expired(ctrl, spins))) break;
                if (expired) break;
            }
        }
        lb_store(&status[(size_t)tile * 256 + tid], lb_word(epoch, LB_P, excl + total));
    }
    // digit tid's keys go to [goff0 + excl, goff0 + excl + total): checked
    // once per digit here instead of once per key below; a tile with an
    // inconsistent offset (a corrupted chain) writes nothing and reports.
    const bool bad = (uint64_t)goff0[tid] + excl + total > n;
    L.goff[tid] = goff0[tid] + excl - lstart;
    if (block_barrier_or(bad)) {
        if (tid == 0) __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }

    // reorder the tile by digit in LDS, then write out coalesced
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (valid(j)) s_keys[L.cnt[w][xf.digit(k[j], shift)] + r[j]] = k[j];
    }
    block_barrier();
    const uint32_t tile_n = FULL ? (uint32_t)TILE : n - base;
    uint32_t od[VB > 0 ? IPT : 1];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const uint32_t i = j * kSortThreads + tid;
        if (FULL || i < tile_n) {
            const KB key = s_keys[i];
            const uint32_t dg = xf.digit(key, shift);
            if constexpr (VB > 0) od[j] = dg;
            kout[L.goff[dg] + i] = key;
        }
    }
    if constexpr (VB > 0) {
        block_barrier();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (valid(j)) s_vals[L.cnt[w][xf.digit(k[j], shift)] + r[j]] = v[j];
        }
        block_barrier();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t i = j * kSortThreads + tid;
            if (FULL || i < tile_n) vout[L.goff[od[j]] + i] = s_vals[i];
        }
    }
}

template <typename KB, int VB, int IPT, int RANK, typename XF>
BDEV void onesweep_tile(const KB *__restrict__ kin, KB *__restrict__ kout,
                        const typename ValT<VB>::type *__restrict__ vin, typename ValT<VB>::type *__restrict__ vout,
                        uint32_t n, uint32_t tile, uint32_t shift, const uint32_t *__restrict__ goff0, XF xf,
                        Ctrl *ctrl, uint64_t *__restrict__ status, uint32_t back)
{
    __shared__ OnesweepLds<KB, VB, IPT> lds;
    if ((uint64_t)(tile + 1) * (kSortThreads * IPT) <= n)
        onesweep_body<KB, VB, IPT, RANK, XF, true>(lds, kin, kout, vin, vout, n, tile, shift, goff0, xf, ctrl, status,
                                                   back);
    else
        onesweep_body<KB, VB, IPT, RANK, XF, false>(lds, kin, kout, vin, vout, n, tile, shift, goff0, xf, ctrl,
                                                    status, back);
}

// Four workgroups per CU need <= 128 VGPRs per lane (4 waves per SIMD): the
// keys-only tiles and the 4-byte-value tiles up to 24 keys per thread meet
// that without spilling (checked in the ISA listing); larger tiles keep the
// compiler's choice.
template <int VB, int IPT>
constexpr int onesweep_waves_per_eu() { return (VB == 0 || (VB <= 4 && IPT <= 24)) ? 4 : 1; }

}  // namespace bcmi
