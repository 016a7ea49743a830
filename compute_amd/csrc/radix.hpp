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
// x of the key:   key = (neg ? -x : x) ^ (flt ? (-(x >> msb) | signbit) : cx)
//   ascending:  unsigned cx=0; signed cx=signbit; float flt
//   descending: unsigned cx=~0 ((T)(-1)-x == ~x); signed neg, cx=signbit;
//               float neg, flt
template <typename KB>
struct KeyXf {
    uint32_t neg;
    uint32_t flt;
    KB cx;
};

template <typename KB>
BDEV KB radix_key(KB x, const KeyXf<KB> &f)
{
    constexpr int W = sizeof(KB) * 8;
    using SK = std::make_signed_t<KB>;
    KB t = f.neg ? (KB)(KB(0) - x) : x;
    KB sgn = (KB)(KB(1) << (W - 1));
    KB fm = f.flt ? (KB)((KB)((SK)x >> (W - 1)) | sgn) : f.cx;
    return (KB)(t ^ fm);
}

template <typename KB>
BDEV uint32_t digit_of(KB x, const KeyXf<KB> &f, uint32_t shift)
{
    return (uint32_t)(radix_key(x, f) >> shift) & 0xFFu;
}

template <int VB>
struct ValT;
template <> struct ValT<0> { using type = uint8_t; };
template <> struct ValT<1> { using type = uint8_t; };
template <> struct ValT<2> { using type = uint16_t; };
template <> struct ValT<4> { using type = uint32_t; };
template <> struct ValT<8> { using type = uint64_t; };
struct alignas(16) v16_t { uint64_t a, b; };
template <> struct ValT<16> { using type = v16_t; };

// Rank the IPT keys a lane holds (wave-striped: slot j of lane l is input
// position j*64 + l of the wave's segment) within the wave: r[j] = number of
// earlier keys of the wave with the same digit.  cnt/peer are the wave's 256
// counters (zeroed by the caller) and 256 peer-mask slots (left zeroed).
// Each lane ORs its bit into its digit's slot (ds_or_b64), reads the slot
// back as the peer mask; the lowest peer advances the counter and clears the
// slot.  valid(j) excludes padding slots of a partial tile.
template <int IPT, typename Valid>
BDEV void rank_in_wave(const uint32_t (&d)[IPT], uint32_t (&r)[IPT], uint32_t *cnt, uint64_t *peer, Valid valid)
{
    const uint64_t lanebit = 1ull << (threadIdx.x & 63);
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool ok = valid(j);
        if (ok) __hip_atomic_fetch_or(&peer[d[j]], lanebit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint64_t peers = __hip_atomic_load(&peer[d[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t below = mbcnt(peers);
        const uint32_t old = cnt[d[j]];
        r[j] = old + below;
        if (ok && below == 0) {
            cnt[d[j]] = old + popc64(peers);
            __hip_atomic_store(&peer[d[j]], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

// Block-wide exclusive scan of one value per thread (256 threads).
BDEV uint32_t block_exclusive_sum256(uint32_t v, uint32_t *s_wsum)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t pre = inc - v;
    for (uint32_t i = 0; i < w; ++i) pre += s_wsum[i];
    return pre;
}

// ------------------------------------------------------------ onesweep pass
// Ranking within a wave (RANK selects the method):
//   kRankLdsMask: each lane ORs its lane bit into a per-wave 64-bit LDS slot
//     for its digit (ds_or_b64) and reads the slot back: the peer mask in
//     ~10 VALU ops + 5 LDS ops per key;
//   kRankBallot: 8 ballots, one per digit bit, AND-ed per lane (~50 VALU
//     ops per key).
// Either way: rank = wave digit counter + peers below this lane, and the
// lowest peer advances the counter; keys of one wave are ranked slot by slot
// (j) in lane order, which is their input order, so the sort is stable.
enum { kRankLdsMask = 0, kRankBallot = 1 };

// One onesweep tile: keys [tile*TILE, tile*TILE + TILE) of an n-key chain.
// status points at the chain's first tile; goff0 holds the chain's 256
// exclusive digit offsets for this pass.
template <typename KB, int VB, int IPT, int RANK>
BDEV void onesweep_tile(const KB *__restrict__ kin, KB *__restrict__ kout,
                        const typename ValT<VB>::type *__restrict__ vin, typename ValT<VB>::type *__restrict__ vout,
                        uint32_t n, uint32_t tile, uint32_t shift, const uint32_t *__restrict__ goff0,
                        KeyXf<KB> xf, Ctrl *ctrl, uint64_t *__restrict__ status, uint32_t epoch)
{
    using V = typename ValT<VB>::type;
    constexpr int TILE = kSortThreads * IPT;
    constexpr int SB = (int)sizeof(KB) > VB ? (int)sizeof(KB) : VB;
    __shared__ uint32_t s_cnt[4][256];
    __shared__ uint64_t s_peer[4][256];
    __shared__ uint32_t s_goff[256];
    __shared__ uint32_t s_wsum[4];
    __shared__ __attribute__((aligned(16))) unsigned char s_stage[TILE * SB];
    KB *s_keys = (KB *)s_stage;
    V *s_vals = (V *)s_stage;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int i = tid; i < 4 * 256; i += kSortThreads) {
        (&s_cnt[0][0])[i] = 0;
        (&s_peer[0][0])[i] = 0;
    }
    __syncthreads();
    const uint32_t base = tile * (uint32_t)TILE;
    const bool full = (uint64_t)base + TILE <= n;
    const uint32_t wb = base + w * 64 * IPT + lane;

    // wave-striped load: lane l of wave w holds tile positions w*64*IPT + j*64 + l
    KB k[IPT];
    V v[IPT];
    uint32_t d[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        uint32_t idx = wb + j * 64;
        bool ok = full || idx < n;
        k[j] = ok ? kin[idx] : KB(0);
        if constexpr (VB > 0) {
            if (ok) v[j] = vin[idx];
        }
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) d[j] = digit_of(k[j], xf, shift);

    // rank within the wave: peers = lanes whose digit matches; rank = wave
    // counter + peers below; the lowest peer bumps the counter.
    uint32_t r[IPT];
    const uint64_t lanebit = 1ull << lane;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool ok = full || (wb + j * 64 < n);
        uint64_t peers;
        if constexpr (RANK == kRankLdsMask) {
            if (ok) __hip_atomic_fetch_or(&s_peer[w][d[j]], lanebit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            peers = __hip_atomic_load(&s_peer[w][d[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            peers = full ? ~0ull : ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                uint32_t nb = ((d[j] >> b) & 1u) - 1u;  // 0 if the bit is set, ~0 if clear
                uint64_t m = ballot(nb == 0u);
                peers &= m ^ (((uint64_t)nb << 32) | nb);
            }
        }
        uint32_t below = mbcnt(peers);
        uint32_t old = s_cnt[w][d[j]];
        r[j] = old + below;
        if (ok && below == 0) {
            s_cnt[w][d[j]] = old + popc64(peers);
            if constexpr (RANK == kRankLdsMask)
                __hip_atomic_store(&s_peer[w][d[j]], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();

    // per digit (thread tid == digit): wave prefixes and the tile count
    const uint32_t c0 = s_cnt[0][tid], c1 = s_cnt[1][tid], c2 = s_cnt[2][tid], c3 = s_cnt[3][tid];
    const uint32_t total = c0 + c1 + c2 + c3;
    if (tile == 0) lb_store(&status[tid], lb_word(epoch, LB_P, total));
    else lb_store(&status[(size_t)tile * 256 + tid], lb_word(epoch, LB_A, total));
    uint32_t inc = wave_inclusive_scan(total, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t lstart = inc - total;
    for (uint32_t i = 0; i < w; ++i) lstart += s_wsum[i];
    s_cnt[0][tid] = lstart;
    s_cnt[1][tid] = lstart + c0;
    s_cnt[2][tid] = lstart + c0 + c1;
    s_cnt[3][tid] = lstart + c0 + c1 + c2;
    // decoupled look-back for this digit: fetch kLookback predecessors'
    // status words at once (one round trip), fold them in tile order until
    // an inclusive prefix; stop and re-poll at a word not yet published.
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
            if (stalled && used == 0 && spin_expired(ctrl, spins)) break;
        }
        lb_store(&status[(size_t)tile * 256 + tid], lb_word(epoch, LB_P, excl + total));
    }
    s_goff[tid] = goff0[tid] + excl - lstart;
    __syncthreads();

    // reorder the tile by digit in LDS, then write out coalesced
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (full || wb + j * 64 < n) s_keys[s_cnt[w][d[j]] + r[j]] = k[j];
    }
    __syncthreads();
    const uint32_t tile_n = full ? (uint32_t)TILE : n - base;
    uint32_t od[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        uint32_t i = j * kSortThreads + tid;
        if (i < tile_n) {
            KB key = s_keys[i];
            od[j] = digit_of(key, xf, shift);
            uint32_t dst = s_goff[od[j]] + i;
            if (dst < n) kout[dst] = key;
            else __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if constexpr (VB > 0) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (full || wb + j * 64 < n) s_vals[s_cnt[w][d[j]] + r[j]] = v[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            uint32_t i = j * kSortThreads + tid;
            if (i < tile_n) {
                uint32_t dst = s_goff[od[j]] + i;
                if (dst < n) vout[dst] = s_vals[i];
            }
        }
    }
}

template <typename KB, int VB, int IPT, int RANK = kRankLdsMask>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(const KB *__restrict__ kin, KB *__restrict__ kout,
                                                           const typename ValT<VB>::type *__restrict__ vin,
                                                           typename ValT<VB>::type *__restrict__ vout, uint32_t n,
                                                           uint32_t shift, uint32_t pass, KeyXf<KB> xf, Ctrl *ctrl,
                                                           uint64_t *__restrict__ status, uint32_t back)
{
    onesweep_tile<KB, VB, IPT, RANK>(kin, kout, vin, vout, n, blockIdx.x, shift, ctrl->offsets[pass], xf, ctrl,
                                     status, launch_epoch(ctrl, back));
}

}  // namespace bcmi
