// device.hpp -- gfx950 device primitives shared by the kernels.
//
// Wave64 building blocks (DPP row shifts/broadcasts, ballots, mbcnt),
// decoupled look-back status words, the last-block protocol that keeps the
// queue's control block zeroed, and error-free float transforms.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "runtime.hpp"

namespace bcmi {

#define BDEV __device__ __forceinline__
#define BHD __host__ __device__ __forceinline__

// Work-group barrier that first waits for the wave's own LDS operations.
// __syncthreads() alone lets a wave pass the barrier with LDS reads still
// outstanding (the compiler waits for them only where it uses the data),
// so another wave's LDS writes after the barrier can land first; with
// several work-groups resident per CU that returned overwritten data.  Every
// kernel here reuses LDS slots across barriers, so every barrier is this one.
BDEV void block_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __syncthreads();
}

BDEV int block_barrier_or(int pred)
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    return __syncthreads_or(pred);
}

constexpr int kWave = 64;
constexpr int kLdsBytesPerCU = 160 * 1024;  // gfx950 LDS per compute unit

BDEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

BDEV uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of m in lanes below this lane
BDEV uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

BDEV uint32_t popc64(uint64_t m) { return (uint32_t)__builtin_popcountll(m); }

// ------------------------------------------------------------------ DPP
// dpp_ctrl encodings (GFX9): row_shr:n = 0x110+n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143, wave_shr:1 = 0x138.  Lanes whose source is invalid
// or whose row is masked off keep `old`.
template <int CTRL, int ROWMASK = 0xf, int BANKMASK = 0xf>
BDEV uint32_t dpp32(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWMASK, BANKMASK, false);
}

// Any trivially copyable value whose size is a multiple of 4 bytes moves
// as that many 32-bit DPP moves.
template <int N>
struct words {
    uint32_t w[N];
};

template <int CTRL, int ROWMASK = 0xf, int BANKMASK = 0xf, typename T>
BDEV T dpp(T old, T v)
{
    static_assert(sizeof(T) % 4 == 0, "dpp on multiples of 4 bytes");
    constexpr int N = sizeof(T) / 4;
    words<N> o = __builtin_bit_cast(words<N>, old), x = __builtin_bit_cast(words<N>, v), r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.w[i] = dpp32<CTRL, ROWMASK, BANKMASK>(o.w[i], x.w[i]);
    return __builtin_bit_cast(T, r);
}

template <typename T>
BDEV T readlane(T v, int lane)
{
    static_assert(sizeof(T) % 4 == 0, "readlane on multiples of 4 bytes");
    constexpr int N = sizeof(T) / 4;
    words<N> x = __builtin_bit_cast(words<N>, v), r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.w[i] = (uint32_t)__builtin_amdgcn_readlane((int)x.w[i], lane);
    return __builtin_bit_cast(T, r);
}

// Inclusive scan across the 64 lanes: Hillis-Steele inside each 16-lane
// row (row_shr 1,2,4,8), then row_bcast:15 into rows 1,3 and row_bcast:31
// into rows 2,3.  `op` must be associative; `id` its identity.
template <typename T, typename Op>
BDEV T wave_inclusive_scan(T x, T id, Op op)
{
    x = op(dpp<0x111>(id, x), x);
    x = op(dpp<0x112>(id, x), x);
    x = op(dpp<0x114>(id, x), x);
    x = op(dpp<0x118>(id, x), x);
    x = op(dpp<0x142, 0xa>(id, x), x);
    x = op(dpp<0x143, 0xc>(id, x), x);
    return x;
}

// inclusive -> exclusive: shift the whole wave right by one lane
template <typename T>
BDEV T wave_shr1(T incl, T id) { return dpp<0x138>(id, incl); }

template <typename T, typename Op>
BDEV T wave_reduce(T x, T id, Op op) { return readlane(wave_inclusive_scan(x, id, op), 63); }

// --------------------------------------------------- decoupled look-back
// A status word is [tag:32 | payload:32]; tag = epoch << 2 | flag.  A tag
// whose epoch differs from the current launch's reads as "not published",
// so the array never needs clearing between launches.  Payloads wider than
// 32 bits are split over several words that carry the same tag; a reader
// accepts them only when all tags agree.
enum : uint32_t { LB_X = 0, LB_A = 1, LB_P = 2 };

BDEV uint64_t lb_word(uint32_t epoch, uint32_t flag, uint32_t payload)
{
    return ((uint64_t)((epoch << 2) | flag) << 32) | payload;
}
BDEV uint32_t lb_tag(uint64_t w) { return (uint32_t)(w >> 32); }
BDEV uint32_t lb_flag(uint64_t w) { return (uint32_t)(w >> 32) & 3u; }
BDEV uint32_t lb_epoch(uint64_t w) { return (uint32_t)(w >> 34); }
BDEV uint32_t lb_payload(uint64_t w) { return (uint32_t)w; }

BDEV void lb_store(uint64_t *p, uint64_t w) { __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
BDEV uint64_t lb_load(const uint64_t *p)
{
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Epoch of a look-back launch.  Before a sequence of L look-back launches
// the host enqueues k_epoch_advance(L) (one thread: ctrl->epoch += L); launch
// j of the sequence passes back = L - j and tags its words with
// ((epoch - back) mod (2^30 - 1)) + 1, so every launch on the queue gets a
// distinct non-zero epoch, graph replays included.  No per-block atomic is
// involved: device-scope atomics on one address serialise across XCDs.
//
// The epoch is a plain load at kernel entry.  ctrl->epoch is written
// only by k_epoch_advance, an earlier kernel on the same stream, so it is
// constant for this launch and visible to an ordinary load; with a uniform
// address that load can be served by the scalar cache, off the vector
// memory path the tile's data uses.
BDEV uint32_t launch_epoch_plain(const Ctrl *ctrl, uint32_t back)
{
    const uint32_t e = ctrl->epoch;
    return ((e - back) % kEpochMask) + 1u;
}

// Tiles are numbered by blockIdx.x.  Workgroups are dispatched in id order,
// so when a block runs, every block with a smaller id has been dispatched
// and is resident or finished: the look-back wait always makes progress.

// Every look-back wait is bounded: after kSpinLimit polls (each a device-
// coherent load plus s_sleep 1; on the order of a second in total), or
// ctrl->spin_limit polls where a test has set one, the wait gives up,
// records ctrl->fault and proceeds, so a protocol error can never hang the
// device.
constexpr uint32_t kSpinLimit = 1u << 20;

BDEV bool spin_expired(Ctrl *ctrl, uint32_t &spins)
{
    const uint32_t limit = ctrl->spin_limit ? ctrl->spin_limit : kSpinLimit;
    if (++spins < limit) {
        __builtin_amdgcn_s_sleep(1);
        return false;
    }
    __hip_atomic_store(&ctrl->fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// ------------------------------------------------ error-free transforms
// Knuth's TwoSum: s + e == a + b exactly (no FMA involved, so contraction
// cannot alter it).
BDEV void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// compensated FP32 accumulator: value = s + c
struct cf32 {
    float s, c;
};

BDEV cf32 cf_add(cf32 a, float x)
{
    float s, e;
    two_sum(a.s, x, s, e);
    return cf32{s, a.c + e};
}

BDEV cf32 cf_merge(cf32 a, cf32 b)
{
    float s, e;
    two_sum(a.s, b.s, s, e);
    return cf32{s, (a.c + b.c) + e};
}

BDEV double cf_value(cf32 a) { return (double)a.s + (double)a.c; }

}  // namespace bcmi

namespace bcmi {

// ------------------------------------------- striped <-> blocked transposes
// A wave-striped tile (chunk q of lane l = wave positions q*256 + l*4 .. +3,
// loaded and stored coalesced) transposed through LDS so that lane l holds
// E = Q*4 consecutive positions l*E .. l*E+E-1, and back.  The buffer has
// one row of E words per lane, padded by 4 words so that the 16-byte
// accesses of both directions spread over the banks.  Each wave has a
// buffer of its own, so a transpose synchronises the wave only: one wave's
// LDS operations execute in order, and wave_lds_sync keeps the compiler from
// moving the reads above the writes.
template <int Q, int NW = 4>
struct BlkBuf {
    static constexpr int E = Q * 4;
    static constexpr int ROW = E + 4;
    uint32_t w[NW][64 * ROW];
};

BDEV void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// striped chunks of one wave -> this lane's row
template <int Q, typename T>
BDEV void to_blocked(const T (&s)[Q][4], T (&b)[Q * 4], uint32_t *buf, uint32_t lane)
{
    constexpr int E = Q * 4, ROW = E + 4;
    static_assert(sizeof(T) == 4, "4-byte elements");
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int p = q * 256 + (int)lane * 4;  // first of the chunk's 4 wave positions
        uint4 c;
        c.x = __builtin_bit_cast(uint32_t, s[q][0]);
        c.y = __builtin_bit_cast(uint32_t, s[q][1]);
        c.z = __builtin_bit_cast(uint32_t, s[q][2]);
        c.w = __builtin_bit_cast(uint32_t, s[q][3]);
        *(uint4 *)(buf + (p / E) * ROW + p % E) = c;
    }
    wave_lds_sync();
#pragma unroll
    for (int e = 0; e < E; e += 4) {
        const uint4 c = *(const uint4 *)(buf + lane * ROW + e);
        b[e] = __builtin_bit_cast(T, c.x);
        b[e + 1] = __builtin_bit_cast(T, c.y);
        b[e + 2] = __builtin_bit_cast(T, c.z);
        b[e + 3] = __builtin_bit_cast(T, c.w);
    }
    wave_lds_sync();
}

// this lane's row -> striped chunks of one wave
template <int Q, typename T>
BDEV void to_striped(const T (&b)[Q * 4], T (&s)[Q][4], uint32_t *buf, uint32_t lane)
{
    constexpr int E = Q * 4, ROW = E + 4;
#pragma unroll
    for (int e = 0; e < E; e += 4) {
        uint4 c;
        c.x = __builtin_bit_cast(uint32_t, b[e]);
        c.y = __builtin_bit_cast(uint32_t, b[e + 1]);
        c.z = __builtin_bit_cast(uint32_t, b[e + 2]);
        c.w = __builtin_bit_cast(uint32_t, b[e + 3]);
        *(uint4 *)(buf + lane * ROW + e) = c;
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int p = q * 256 + (int)lane * 4;
        const uint4 c = *(const uint4 *)(buf + (p / E) * ROW + p % E);
        s[q][0] = __builtin_bit_cast(T, c.x);
        s[q][1] = __builtin_bit_cast(T, c.y);
        s[q][2] = __builtin_bit_cast(T, c.z);
        s[q][3] = __builtin_bit_cast(T, c.w);
    }
    wave_lds_sync();
}

// ------------------------------------------- wave-parallel decoupled look-back
// A carry of 4*N bytes occupies N consecutive status words per tile.
template <typename Cr>
BDEV void lb_publish(uint64_t *status, uint32_t tile, uint32_t epoch, uint32_t flag, Cr v)
{
    static_assert(sizeof(Cr) % 4 == 0, "carry of a multiple of 4 bytes");
    constexpr int N = sizeof(Cr) / 4;
    words<N> x = __builtin_bit_cast(words<N>, v);
#pragma unroll
    for (int i = 0; i < N; ++i) lb_store(&status[(size_t)N * tile + i], lb_word(epoch, flag, x.w[i]));
}

template <typename Cr>
BDEV uint32_t lb_read(const uint64_t *status, size_t tile, uint32_t epoch, Cr &v)
{
    constexpr int N = sizeof(Cr) / 4;
    uint64_t w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = lb_load(&status[(size_t)N * tile + i]);
    bool ok = lb_epoch(w[0]) == epoch;
#pragma unroll
    for (int i = 1; i < N; ++i) ok = ok && (lb_tag(w[i]) == lb_tag(w[0]));
    if (!ok) return LB_X;
    words<N> x;
#pragma unroll
    for (int i = 0; i < N; ++i) x.w[i] = lb_payload(w[i]);
    v = __builtin_bit_cast(Cr, x);
    return lb_flag(w[0]);
}

// Called by all 64 lanes of one wave of a tile > 0: folds the predecessors'
// aggregates / inclusive prefixes 64 tiles at a time until an inclusive
// prefix is found and returns the tile's exclusive prefix (without init;
// tile 0's inclusive prefix carries it).  Publishes nothing, so it can run
// before the tile's own aggregate is known.  `cop(a, b)` must be
// associative; a is the earlier operand.  Lane l reads tile (base - l), so
// the lanes are folded with the operands swapped to keep tile order.  When
// a word before the first inclusive prefix is not yet published, only that
// one word is polled (by lane 0) until it appears, and then the 64-wide
// read is repeated: a thousand waiting tiles re-reading 64 words each would
// load the memory system with coherent reads.
template <typename Cr, typename Op>
BDEV Cr lookback_prefix(const uint64_t *status, uint32_t tile, uint32_t epoch, Cr cid, Op cop, Ctrl *ctrl)
{
    uint32_t spins = 0;
    const uint32_t lane = lane_id();
    Cr acc = cid;
    int64_t base = (int64_t)tile - 1;
    bool gave_up = false;
    while (true) {
        const int64_t idx = base - (int64_t)lane;
        Cr v = cid;
        uint32_t flag = LB_P;  // before tile 0: the identity, as an inclusive prefix
        if (idx >= 0) flag = lb_read(status, (size_t)idx, epoch, v);
        const uint64_t pm = ballot(flag == LB_P);
        const uint64_t xm = ballot(flag == LB_X);
        const uint32_t first = pm ? (uint32_t)__builtin_ctzll(pm) : 64u;
        const uint64_t rel = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
        if ((xm & rel) && !gave_up) {
            const int64_t widx = base - (int64_t)__builtin_ctzll(xm & rel);
            while (true) {
                uint32_t f = LB_X;
                if (lane == 0) {
                    Cr t;
                    f = lb_read(status, (size_t)widx, epoch, t);
                }
                if (__builtin_amdgcn_readfirstlane(f) != LB_X) break;
                if (spin_expired(ctrl, spins)) {
                    gave_up = true;  // recorded as a fault; finish with what is there
                    break;
                }
            }
            continue;
        }
        Cr mine = lane <= first ? v : cid;
        Cr older = wave_reduce(mine, cid, [&](Cr a, Cr b) { return cop(b, a); });
        acc = cop(older, acc);
        if (first < 64) break;
        base -= 64;
    }
    return acc;
}

// Called by all 64 lanes of one wave.  Publishes the tile's aggregate,
// obtains the exclusive prefix by look-back, publishes the tile's inclusive
// prefix and returns its exclusive prefix (init folded in at tile 0).
template <typename Cr, typename Op>
BDEV Cr lookback_scan(uint64_t *status, uint32_t tile, uint32_t epoch, Cr agg, Cr init, Cr cid, Op cop,
                      Ctrl *ctrl)
{
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) lb_publish(status, 0, epoch, LB_P, cop(init, agg));
        return init;
    }
    if (lane == 0) lb_publish(status, tile, epoch, LB_A, agg);
    const Cr acc = lookback_prefix<Cr>(status, tile, epoch, cid, cop, ctrl);
    if (lane == 0) lb_publish(status, tile, epoch, LB_P, cop(acc, agg));
    return acc;
}

}  // namespace bcmi
