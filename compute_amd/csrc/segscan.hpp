// segscan.hpp -- reduce_by_key and the segmented (by-key) scans.
//
// Reference (GPU column): reduce_by_key_with_scan
// (algorithm/detail/reduce_by_key_with_scan.hpp:429-488): head flags +
// inclusive_scan for the output index (:52-98), a segmented Hillis-Steele
// scan per work-group writing carry-outs (:104-179), a one-work-group
// carry-in scan (:183-293), a final segmented scan that adds the carry-in
// and writes the run's last key and value at its index (:301-400), then a
// blocking read of the run count (:485-487).  About 40 bytes per element
// over 5+ launches.
//
// Here: one single-pass launch.  Each element contributes the pair
// (is_head, value); pairs combine with the segmented operator
//   (n1, v1) . (n2, v2) = (n1 + n2, n2 ? v2 : v1 op v2)
// (associative, not commutative).  Tiles scan pairs in registers (chunk
// folds, DPP wave scans, an LDS step over the 4 waves) and get their
// exclusive prefix pair by ordered decoupled look-back.  At each run head
// i > 0 the exclusive pair at i is the finished run i-1: its run index is
// (heads before i) - 1, and the key written is key[i-1] -- the run's last
// key, as the reference writes (reduce_by_key_with_scan.hpp:381-384,
// serial_reduce_by_key.hpp:80,87).  Float plus keeps compensated FP32
// pairs inside the tile and an FP64 value in the carried pair.
#pragma once

#include "device.hpp"
#include "ops.hpp"

#include <cstring>

namespace bcmi {

constexpr int kSegThreads = 256;

template <typename A>
struct Seg {
    A v;
    uint32_t n;
};

template <typename VT>
using widen_t = std::conditional_t<(sizeof(VT) < 4), std::conditional_t<std::is_signed<VT>::value, int32_t, uint32_t>, VT>;

template <typename VT, int OP>
struct SegPlain {
    using A = widen_t<VT>;
    using CA = A;
    static BHD A id() { return op_identity<A, OP>(); }
    static BDEV A add(A a, VT x) { return apply_op<A, OP>(a, (A)x); }
    static BDEV A merge(A a, A b) { return apply_op<A, OP>(a, b); }
    static BDEV CA carry(A a) { return a; }
    static BHD CA cid() { return id(); }
    static BDEV CA cmerge(CA a, CA b) { return apply_op<A, OP>(a, b); }
    // value of (prefix . local): the prefix value counts only if no head
    // occurred in the local part
    static BDEV VT fin(CA p, A s, bool use_p) { return (VT)(use_p ? apply_op<A, OP>(p, s) : s); }
    static BDEV VT fin_local(A s) { return (VT)s; }
    static BDEV VT fin_init(VT init, CA p, A s, bool use_p)
    {
        return (VT)apply_op<A, OP>((A)init, use_p ? apply_op<A, OP>(p, s) : s);
    }
};

struct SegCompF32 {
    using A = cf32;
    using CA = double;
    static BHD A id() { return cf32{0.f, 0.f}; }
    static BDEV A add(A a, float x) { return cf_add(a, x); }
    static BDEV A merge(A a, A b) { return cf_merge(a, b); }
    static BDEV CA carry(A a) { return cf_value(a); }
    static BHD CA cid() { return 0.0; }
    static BDEV CA cmerge(CA a, CA b) { return a + b; }
    static BDEV float fin(CA p, A s, bool use_p) { return (float)(use_p ? p + cf_value(s) : cf_value(s)); }
    static BDEV float fin_local(A s) { return (float)cf_value(s); }
    static BDEV float fin_init(float init, CA p, A s, bool use_p)
    {
        return (float)((double)init + (use_p ? p + cf_value(s) : cf_value(s)));
    }
};

// The same precision class with FP64 accumulators: float elements widened
// once, run sums and carries in double (one add per element instead of a
// TwoSum pair's six instructions; FP64 vector math is half the FP32 rate).
struct SegF64AccF32 {
    using A = double;
    using CA = double;
    static BHD A id() { return 0.0; }
    static BDEV A add(A a, float x) { return a + (double)x; }
    static BDEV A merge(A a, A b) { return a + b; }
    static BDEV CA carry(A a) { return a; }
    static BHD CA cid() { return 0.0; }
    static BDEV CA cmerge(CA a, CA b) { return a + b; }
    static BDEV float fin(CA p, A s, bool use_p) { return (float)(use_p ? p + s : s); }
    static BDEV float fin_local(A s) { return (float)s; }
    static BDEV float fin_init(float init, CA p, A s, bool use_p)
    {
        return (float)((double)init + (use_p ? p + s : s));
    }
};

template <typename A, typename P>
BDEV Seg<A> seg_merge(Seg<A> a, Seg<A> b)
{
    return Seg<A>{b.n ? b.v : P::merge(a.v, b.v), a.n + b.n};
}

template <typename KT>
BDEV bool key_eq(KT a, KT b) { return a == b; }  // equal_to<T>: -0 == +0, NaN != NaN

enum { SEG_RBK = 0, SEG_INCL = 1, SEG_EXCL = 2 };

template <typename P>
struct SegLds {
    Seg<typename P::A> wave[4];
    Seg<typename P::CA> prefix;
};

// V consecutive keys, read or written as one access (16 or 32 bytes here;
// 8 when 8-byte values pair with 4-byte keys)
template <typename KT, int V>
struct alignas(V * sizeof(KT) >= 16 ? 16 : V * sizeof(KT)) key_chunk {
    KT v[V];
};

// One tile: Q 16-byte value chunks per thread, wave-striped as in the scan
// (chunk q of lane l of wave w covers tile elements w*WSEG + q*64*V + l*V ..
// +V-1).  FULL: the tile lies below n.  VEC: keys and values (and, for the
// scans, the output) are 16-byte aligned, so full chunks move as vectors.
// The key before each chunk comes from the lane below (DPP), from lane 63
// of the previous chunk, or, for the wave's first chunk, from memory.
template <typename KT, typename VT, typename P, int MODE, int Q, bool VEC, bool FULL>
BDEV void segscan_body(SegLds<P> &L, const KT *__restrict__ keys, const VT *vals, VT *out,
                       KT *__restrict__ keys_out, size_t n, Ctrl *ctrl, uint64_t *status, VT init,
                       uint64_t *count_out, uint32_t epoch)
{
    using A = typename P::A;
    using CA = typename P::CA;
    using S = Seg<A>;
    using C = Seg<CA>;
    constexpr int V = 16 / sizeof(VT);
    constexpr int WSEG = Q * 64 * V;
    constexpr int TILE = 4 * WSEG;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const size_t wstart = (size_t)tile * TILE + (size_t)__builtin_amdgcn_readfirstlane(w) * WSEG;
    const KT *wk = keys + wstart;
    const VT *wv = vals + wstart;
    const int64_t rem = FULL ? (int64_t)WSEG : (int64_t)n - (int64_t)wstart;
    auto ok = [&](int q, int k) { return FULL || (int64_t)(q * 64 * V + lane * V + k) < rem; };

    VT x[Q][V];
    KT kc[Q][V];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int o = q * 64 * V + lane * V;
        if (VEC && FULL) {
            const key_chunk<KT, V> kq = *(const key_chunk<KT, V> *)(wk + o);
            const key_chunk<VT, V> vq = *(const key_chunk<VT, V> *)(wv + o);
#pragma unroll
            for (int k = 0; k < V; ++k) {
                kc[q][k] = kq.v[k];
                x[q][k] = vq.v[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < V; ++k) {
                kc[q][k] = ok(q, k) ? wk[o + k] : KT(0);
                x[q][k] = ok(q, k) ? wv[o + k] : VT(0);
            }
        }
    }
    // head flags: element i starts a run iff i == 0 or !(key[i-1] == key[i])
    const KT kfirst = (wstart > 0 && wstart - 1 < n) ? keys[wstart - 1] : KT(0);
    uint32_t headm[Q];
    KT kprev[Q];  // the key before each chunk
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        KT prev = wave_shr1(kc[q][V - 1], KT(0));
        if (lane == 0) prev = q == 0 ? kfirst : readlane(kc[q - 1][V - 1], 63);
        kprev[q] = prev;
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const KT before = k == 0 ? prev : kc[q][k - 1];
            if (!key_eq(before, kc[q][k])) m |= 1u << k;
        }
        if (q == 0 && lane == 0 && wstart == 0) m |= 1u;
        headm[q] = m;
    }
    // chunk folds of (head, value) pairs, wave scans carried across q
    const S sid{P::id(), 0u};
    auto mrg = [](S a, S b) { return seg_merge<A, P>(a, b); };
    S carry = sid;
    S pre[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        S t = sid;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (ok(q, k)) {
                if (headm[q] >> k & 1u) t = S{P::add(P::id(), x[q][k]), t.n + 1};
                else t.v = P::add(t.v, x[q][k]);
            }
        }
        const S inc = wave_inclusive_scan(t, sid, mrg);
        pre[q] = mrg(carry, wave_shr1(inc, sid));
        carry = mrg(carry, readlane(inc, 63));
    }
    if (lane == 0) L.wave[w] = carry;
    block_barrier();
    S wpre = sid;
    for (uint32_t i = 0; i < w; ++i) wpre = mrg(wpre, L.wave[i]);
    if (w == 0) {
        const S total = mrg(mrg(L.wave[0], L.wave[1]), mrg(L.wave[2], L.wave[3]));
        const C agg{P::carry(total.v), total.n};
        const C cid{P::cid(), 0u};
        const C prefix = lookback_scan<C>(status, tile, epoch, agg, cid, cid, [](C a, C b) {
            return C{b.n ? b.v : P::cmerge(a.v, b.v), a.n + b.n};
        }, ctrl);
        if (lane == 0) L.prefix = prefix;
    }
    block_barrier();
    const C prefix = L.prefix;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int o = q * 64 * V + lane * V;
        S s = mrg(wpre, pre[q]);  // exclusive local pair before the chunk
        VT y[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (!ok(q, k)) continue;
            const size_t i = wstart + o + k;
            const bool head = headm[q] >> k & 1u;
            if constexpr (MODE == SEG_RBK) {
                if (head && i > 0) {
                    // finished run i-1 = the exclusive pair at i
                    const uint32_t nb = prefix.n + s.n;
                    if (nb >= 1 && nb <= n) {
                        keys_out[nb - 1] = k == 0 ? kprev[q] : kc[q][k - 1];
                        out[nb - 1] = P::fin(prefix.v, s.v, s.n == 0);
                    } else {
                        __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            } else if constexpr (MODE == SEG_EXCL) {
                y[k] = head ? init : P::fin_init(init, prefix.v, s.v, s.n == 0);
            }
            if (head) s = S{P::add(P::id(), x[q][k]), s.n + 1};
            else s.v = P::add(s.v, x[q][k]);
            if constexpr (MODE == SEG_INCL) y[k] = P::fin(prefix.v, s.v, s.n == 0);
            if constexpr (MODE == SEG_RBK) {
                if (i == n - 1) {
                    const uint32_t nb = prefix.n + s.n;
                    if (nb >= 1 && nb <= n) {
                        keys_out[nb - 1] = kc[q][k];
                        out[nb - 1] = P::fin(prefix.v, s.v, s.n == 0);
                    } else {
                        __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    *count_out = nb;
                }
            }
        }
        if constexpr (MODE != SEG_RBK) {
            VT *dst = out + wstart + o;
            if (VEC && FULL) {
                key_chunk<VT, V> c;
#pragma unroll
                for (int k = 0; k < V; ++k) c.v[k] = y[k];
                *(key_chunk<VT, V> *)dst = c;
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k)
                    if (ok(q, k)) dst[k] = y[k];
            }
        }
    }
}

// ---------------------------------------------------------- blocked form
// For 4-byte keys and values the tile is transposed through LDS (to_blocked
// in device.hpp) so that each lane folds E = Q*4 consecutive elements and the
// wave needs one scan of the folds instead of one per chunk: for the
// compensated pairs that is most of the arithmetic.
template <typename P, int Q, int NW>
struct SegLdsBlk {
    Seg<typename P::A> wave[NW];
    Seg<typename P::CA> prefix;
    BlkBuf<Q, NW> t;
};

// a[i] for a lane-varying i < E, without scratch: a select tree on the
// bits of i (E - 1 selects)
template <int E, typename T>
BDEV T pick(const T (&a)[E], uint32_t i)
{
    static_assert(E >= 2 && (E & (E - 1)) == 0, "a power of two");
    T v[E];
#pragma unroll
    for (int j = 0; j < E; ++j) v[j] = a[j];
#pragma unroll
    for (int b = 0, m = E; m > 1; ++b, m /= 2) {
        const bool hi = (i >> b) & 1u;
#pragma unroll
        for (int j = 0; j < m / 2; ++j) v[j] = hi ? v[2 * j + 1] : v[2 * j];
    }
    return v[0];
}

template <typename KT, typename VT, typename P, int MODE, int Q, int NW, bool VEC, bool FULL>
BDEV void segscan_body_blk(SegLdsBlk<P, Q, NW> &L, const KT *__restrict__ keys, const VT *vals, VT *out,
                           KT *__restrict__ keys_out, size_t n, Ctrl *ctrl, uint64_t *status, VT init,
                           uint64_t *count_out, uint32_t epoch)
{
    static_assert(sizeof(KT) == 4 && sizeof(VT) == 4, "blocked form: 4-byte keys and values");
    using A = typename P::A;
    using CA = typename P::CA;
    using S = Seg<A>;
    using C = Seg<CA>;
    constexpr int V = 4;
    constexpr int E = Q * V;  // elements per lane
    constexpr int WSEG = Q * 64 * V;
    constexpr int TILE = NW * WSEG;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const size_t wstart = (size_t)tile * TILE + (size_t)__builtin_amdgcn_readfirstlane(w) * WSEG;
    const KT *wk = keys + wstart;
    const VT *wv = vals + wstart;
    const int64_t rem = FULL ? (int64_t)WSEG : (int64_t)n - (int64_t)wstart;
    auto sok = [&](int q, int k) { return FULL || (int64_t)(q * 64 * V + lane * V + k) < rem; };
    auto ok = [&](int e) { return FULL || (int64_t)(lane * E + e) < rem; };
    uint32_t *buf = L.t.w[w];

    KT kb[E];
    VT xb[E];
    {
        KT kc[Q][V];
        VT x[Q][V];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int o = q * 64 * V + lane * V;
            if (VEC && FULL) {
                const key_chunk<KT, V> kq = *(const key_chunk<KT, V> *)(wk + o);
                const key_chunk<VT, V> vq = *(const key_chunk<VT, V> *)(wv + o);
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    kc[q][k] = kq.v[k];
                    x[q][k] = vq.v[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    kc[q][k] = sok(q, k) ? wk[o + k] : KT(0);
                    x[q][k] = sok(q, k) ? wv[o + k] : VT(0);
                }
            }
        }
        to_blocked<Q>(kc, kb, buf, lane);
        to_blocked<Q>(x, xb, buf, lane);
    }
    // head flags: element i starts a run iff i == 0 or !(key[i-1] == key[i])
    const KT kfirst = (wstart > 0 && wstart - 1 < n) ? keys[wstart - 1] : KT(0);
    KT kprev = wave_shr1(kb[E - 1], KT(0));
    if (lane == 0) kprev = kfirst;
    static_assert(E <= 32, "one head bit per position");
    uint32_t headm = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const KT before = e == 0 ? kprev : kb[e - 1];
        if (!key_eq(before, kb[e])) headm |= 1u << e;
    }
    if (lane == 0 && wstart == 0) headm |= 1u;
    // fold of this lane's (head, value) pairs, one wave scan.  reduce_by_key
    // also keeps, at every head, the run that ends just before it when that
    // run began in this lane (its sum is complete here; parked in the lane's
    // row of the transpose buffer, free now, rather than in registers), and
    // the sum of the lane's elements before its first head (that run came
    // from before the lane and is finished once the lane's prefix is known).
    const S sid{P::id(), 0u};
    auto mrg = [](S a, S b) { return seg_merge<A, P>(a, b); };
    S t = sid;
    VT y[MODE == SEG_RBK ? 1 : E];
    uint32_t *row = buf + lane * BlkBuf<Q, NW>::ROW;
    A tfh = P::id();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (ok(e)) {
            const bool head = headm >> e & 1u;
            if constexpr (MODE == SEG_RBK) {
                if (head) {
                    if (t.n) row[e] = __builtin_bit_cast(uint32_t, P::fin_local(t.v));
                    else tfh = t.v;
                }
            }
            if (head) t = S{P::add(P::id(), xb[e]), t.n + 1};
            else t.v = P::add(t.v, xb[e]);
        }
    }
    const S inc = wave_inclusive_scan(t, sid, mrg);
    const S pre = wave_shr1(inc, sid);
    if (lane == 63) L.wave[w] = inc;
    block_barrier();
    S wpre = sid;
    for (uint32_t i = 0; i < w; ++i) wpre = mrg(wpre, L.wave[i]);
    if (w == 0) {
        S total = L.wave[0];
#pragma unroll
        for (int i = 1; i < NW; ++i) total = mrg(total, L.wave[i]);
        const C agg{P::carry(total.v), total.n};
        const C cid{P::cid(), 0u};
        const C prefix = lookback_scan<C>(status, tile, epoch, agg, cid, cid, [](C a, C b) {
            return C{b.n ? b.v : P::cmerge(a.v, b.v), a.n + b.n};
        }, ctrl);
        if (lane == 0) L.prefix = prefix;
    }
    block_barrier();
    const C prefix = L.prefix;
    const S s0 = mrg(wpre, pre);  // exclusive local pair before this lane's first element
    if constexpr (MODE == SEG_RBK) {
        // runs end before each head (i > 0) and at element n - 1; run r is
        // written at index r, the number of heads before the run's end minus
        // one.  Only the run ending at the lane's first head, and a final
        // run with no head in the lane, continue from before the lane.
        const VT vfirst = P::fin(prefix.v, P::merge(s0.v, tfh), s0.n == 0);
        const uint32_t nb0 = prefix.n + s0.n;  // heads before this lane
        // The lane walks its heads, not its positions: one key and one value
        // store per run end (runs average tens of elements, so a walk over
        // all E positions would issue a mostly idle store pair at each).
        const size_t lbase = wstart + (size_t)lane * E;
        const int64_t lrem = (int64_t)n - (int64_t)lbase;  // positions of the lane below n
        uint32_t hm = lrem <= 0 ? 0u : (lrem >= E ? headm : headm & ((1u << lrem) - 1u));
        if (lbase == 0) hm &= ~1u;  // element 0 ends no run
        while (hm) {
            const int e = __builtin_ctz(hm);
            hm &= hm - 1u;
            const uint32_t hc = __builtin_popcount(headm & ((1u << e) - 1u));  // this lane's heads before e
            const uint32_t nb = nb0 + hc;
            if (nb >= 1 && nb <= n) {
                keys_out[nb - 1] = e == 0 ? kprev : pick(kb, (uint32_t)e - 1u);
                out[nb - 1] = hc == 0 ? vfirst : __builtin_bit_cast(VT, row[e]);
            } else {
                __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lrem >= 1 && lrem <= E) {  // the last run ends at element n - 1, position lrem - 1 here
            const uint32_t e = (uint32_t)lrem - 1u;
            const uint32_t hc = __builtin_popcount(headm & (e == 31 ? ~0u : (2u << e) - 1u));
            const uint32_t nb = nb0 + hc;
            if (nb >= 1 && nb <= n) {
                keys_out[nb - 1] = pick(kb, e);
                out[nb - 1] = hc ? P::fin_local(t.v) : P::fin(prefix.v, P::merge(s0.v, t.v), s0.n == 0);
            } else {
                __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            *count_out = nb;
        }
        return;
    }
    S s = s0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if (!ok(e)) continue;
        const bool head = headm >> e & 1u;
        if constexpr (MODE == SEG_EXCL) y[e] = head ? init : P::fin_init(init, prefix.v, s.v, s.n == 0);
        if (head) s = S{P::add(P::id(), xb[e]), s.n + 1};
        else s.v = P::add(s.v, xb[e]);
        if constexpr (MODE == SEG_INCL) y[e] = P::fin(prefix.v, s.v, s.n == 0);
    }
    if constexpr (MODE != SEG_RBK) {
        VT ys[Q][V];
        to_striped<Q>(y, ys, buf, lane);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int o = q * 64 * V + lane * V;
            VT *dst = out + wstart + o;
            if (VEC && FULL) {
                key_chunk<VT, V> c;
#pragma unroll
                for (int k = 0; k < V; ++k) c.v[k] = ys[q][k];
                *(key_chunk<VT, V> *)dst = c;
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k)
                    if (sok(q, k)) dst[k] = ys[q][k];
            }
        }
    }
}

// NW waves per workgroup: 4, or 8 for the blocked form (a tile twice as big
// at the same elements per lane: half the tiles, so half the look-backs and
// per-tile latencies per element; 16 waves measured slower again).
// Eight-wave workgroups are held to 128 VGPRs (4 waves per SIMD: two
// workgroups per CU, as their LDS allows).
template <typename KT, typename VT, typename P, int MODE, int Q, bool VEC, int NW = 4>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(NW == 8 ? 4 : 1))) void k_segscan(const KT *__restrict__ keys, const VT *vals, VT *out,
                                                     KT *__restrict__ keys_out, size_t n, Ctrl *ctrl,
                                                     uint64_t *status, VT init, uint64_t *count_out,
                                                     uint32_t back)
{
    constexpr size_t TILE = (size_t)NW * Q * 64 * (16 / sizeof(VT));
    const bool full = (size_t)(blockIdx.x + 1) * TILE <= n;
    const uint32_t epoch = launch_epoch_plain(ctrl, back);
    if constexpr (sizeof(KT) == 4 && sizeof(VT) == 4 && Q <= 8) {
        __shared__ SegLdsBlk<P, Q, NW> lds;
        if (full)
            segscan_body_blk<KT, VT, P, MODE, Q, NW, VEC, true>(lds, keys, vals, out, keys_out, n, ctrl, status,
                                                                init, count_out, epoch);
        else
            segscan_body_blk<KT, VT, P, MODE, Q, NW, VEC, false>(lds, keys, vals, out, keys_out, n, ctrl, status,
                                                                 init, count_out, epoch);
    } else {
        static_assert(NW == 4, "the striped form runs 4 waves");
        __shared__ SegLds<P> lds;
        if (full)
            segscan_body<KT, VT, P, MODE, Q, VEC, true>(lds, keys, vals, out, keys_out, n, ctrl, status, init,
                                                        count_out, epoch);
        else
            segscan_body<KT, VT, P, MODE, Q, VEC, false>(lds, keys, vals, out, keys_out, n, ctrl, status, init,
                                                         count_out, epoch);
    }
}

template <typename KT, typename VT, typename P, int MODE, int Q, int NW = 4>
static int launch_seg(bcmi_queue q, const void *keys, const void *vals, void *out, void *keys_out, size_t n,
                      uint64_t *status, VT ini, uint64_t *count_out)
{
    constexpr size_t TILE = (size_t)NW * Q * 64 * (16 / sizeof(VT));
    size_t tiles = (n + TILE - 1) / TILE;
    const bool vec = (((uintptr_t)keys | (uintptr_t)vals | (MODE == SEG_RBK ? 0 : (uintptr_t)out)) & 15) == 0;
    BCMI_TRY(begin_lookback(q, 1));
    if (vec)
        BCMI_LAUNCH(q, MODE == SEG_RBK ? "reduce_by_key" : "segmented_scan",
                    hipLaunchKernelGGL((k_segscan<KT, VT, P, MODE, Q, true, NW>), dim3((unsigned)tiles),
                                       dim3(NW * 64), 0, q->stream, (const KT *)keys, (const VT *)vals, (VT *)out,
                                       (KT *)keys_out, n, q->ctrl, status, ini, count_out, 1u));
    else
        BCMI_LAUNCH(q, MODE == SEG_RBK ? "reduce_by_key" : "segmented_scan",
                    hipLaunchKernelGGL((k_segscan<KT, VT, P, MODE, Q, false, NW>), dim3((unsigned)tiles),
                                       dim3(NW * 64), 0, q->stream, (const KT *)keys, (const VT *)vals, (VT *)out,
                                       (KT *)keys_out, n, q->ctrl, status, ini, count_out, 1u));
    return 0;
}

// Chunks per thread from the tuning cache (key "__bcmi_reduce_by_key" or
// "__bcmi_segscan", param "q": 2, 4 or 8) for 4-byte keys with 4-byte
// values, and waves per workgroup (param "waves": 4, or 8 with q = 8);
// 4 chunks and 4 waves otherwise.
template <typename KT, typename VT, typename P, int MODE>
static int run_seg(bcmi_queue q, const void *keys, const void *vals, void *out, void *keys_out, size_t n,
                   const void *init, uint64_t *count_out)
{
    BCMI_CHECK(n < (1ull << 32), BCMI_INVALID_BUFFER_SIZE, "segmented scan: n=%zu exceeds 2^32-1", n);
    constexpr size_t words = sizeof(Seg<typename P::CA>) / 4;
    const size_t tiles_max = (n + (size_t)4 * 2 * 64 - 1) / ((size_t)4 * 2 * 64) + 1;
    uint64_t *status;
    BCMI_TRY(status_array(q, tiles_max * words, &status));
    VT ini{};
    if (init) memcpy(&ini, init, sizeof(VT));
    if constexpr (sizeof(KT) == 4 && sizeof(VT) == 4) {
        const char *key = MODE == SEG_RBK ? "__bcmi_reduce_by_key" : "__bcmi_segscan";
        const unsigned qq = tune_get(key, "q", 8);
        if (qq == 8 && tune_get(key, "waves", 4) == 8)
            return launch_seg<KT, VT, P, MODE, 8, 8>(q, keys, vals, out, keys_out, n, status, ini, count_out);
        if (qq == 8) return launch_seg<KT, VT, P, MODE, 8>(q, keys, vals, out, keys_out, n, status, ini, count_out);
        if (qq == 2) return launch_seg<KT, VT, P, MODE, 2>(q, keys, vals, out, keys_out, n, status, ini, count_out);
    }
    return launch_seg<KT, VT, P, MODE, 4>(q, keys, vals, out, keys_out, n, status, ini, count_out);
}

template <typename KT, typename VT, int MODE>
static int seg_val_op(bcmi_queue q, int op, int precision, const void *keys, const void *vals, void *out,
                      void *keys_out, size_t n, const void *init, uint64_t *count_out)
{
    if constexpr (std::is_same<VT, float>::value) {
        if (op == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED) {
            // tuning cache key "__bcmi_reduce_by_key" / "__bcmi_segscan",
            // param "acc": 1 = FP64 accumulators, 0 = TwoSum FP32 pairs
            if (tune_get(MODE == SEG_RBK ? "__bcmi_reduce_by_key" : "__bcmi_segscan", "acc", 1))
                return run_seg<KT, VT, SegF64AccF32, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
            return run_seg<KT, VT, SegCompF32, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
        }
    }
    if constexpr (MODE != SEG_RBK) {
        // segmented scans are plus-scans
        return run_seg<KT, VT, SegPlain<VT, BCMI_PLUS>, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
    }
    switch (op) {
    case BCMI_PLUS: return run_seg<KT, VT, SegPlain<VT, BCMI_PLUS>, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_MULTIPLIES:
        return run_seg<KT, VT, SegPlain<VT, BCMI_MULTIPLIES>, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_MIN: return run_seg<KT, VT, SegPlain<VT, BCMI_MIN>, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_MAX: return run_seg<KT, VT, SegPlain<VT, BCMI_MAX>, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
    default: set_error("reduce_by_key/segmented scan: unsupported op %d", op); return BCMI_INVALID_VALUE;
    }
}

template <typename KT, int MODE>
static int seg_vdt(bcmi_queue q, int vdt, int op, int precision, const void *keys, const void *vals, void *out,
                   void *keys_out, size_t n, const void *init, uint64_t *count_out)
{
    switch (vdt) {
    case BCMI_I32: return seg_val_op<KT, int32_t, MODE>(q, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_U32: return seg_val_op<KT, uint32_t, MODE>(q, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_I64: return seg_val_op<KT, int64_t, MODE>(q, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_U64: return seg_val_op<KT, uint64_t, MODE>(q, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_F32: return seg_val_op<KT, float, MODE>(q, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_F64: return seg_val_op<KT, double, MODE>(q, op, precision, keys, vals, out, keys_out, n, init, count_out);
    default: set_error("reduce_by_key/segmented scan: value dtype %d unsupported", vdt); return BCMI_INVALID_VALUE;
    }
}

// One instantiation unit per key type (segscan_<key>.hip) defines these.
#define BCMI_SEG_SIGNATURE(NAME)                                                                   \
    int NAME(bcmi_queue q, int mode, int vdt, int op, int precision, const void *keys, const void *vals, \
             void *out, void *keys_out, size_t n, const void *init, uint64_t *count_out)

BCMI_SEG_SIGNATURE(seg_keys_i32);
BCMI_SEG_SIGNATURE(seg_keys_u32);
BCMI_SEG_SIGNATURE(seg_keys_i64);
BCMI_SEG_SIGNATURE(seg_keys_u64);
BCMI_SEG_SIGNATURE(seg_keys_f32);
BCMI_SEG_SIGNATURE(seg_keys_f64);

#define BCMI_SEG_DEFINE(NAME, KT)                                                                        \
    BCMI_SEG_SIGNATURE(NAME)                                                                             \
    {                                                                                                    \
        switch (mode) {                                                                                  \
        case SEG_RBK: return seg_vdt<KT, SEG_RBK>(q, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out); \
        case SEG_INCL: return seg_vdt<KT, SEG_INCL>(q, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out); \
        default: return seg_vdt<KT, SEG_EXCL>(q, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out); \
        }                                                                                                \
    }

}  // namespace bcmi
