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
    static BDEV float fin_init(float init, CA p, A s, bool use_p)
    {
        return (float)((double)init + (use_p ? p + cf_value(s) : cf_value(s)));
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

// kSegQ = 16-byte value chunks per thread.
template <typename KT, typename VT, typename P, int MODE, int kSegQ>
__global__ __launch_bounds__(kSegThreads) void k_segscan(const KT *__restrict__ keys, const VT *vals, VT *out,
                                                         KT *__restrict__ keys_out, size_t n, Ctrl *ctrl,
                                                         uint64_t *status, VT init, uint64_t *count_out,
                                                         uint32_t back)
{
    using A = typename P::A;
    using CA = typename P::CA;
    using S = Seg<A>;
    using C = Seg<CA>;
    constexpr int V = 16 / sizeof(VT);
    constexpr int WSEG = kSegQ * 64 * V;
    constexpr int TILE = 4 * WSEG;
    __shared__ S s_wave[4];
    __shared__ C s_prefix;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t epoch = launch_epoch(ctrl, back);
    const uint32_t tile = blockIdx.x;
    const size_t tbase = (size_t)tile * TILE;
    const size_t wbase = tbase + (size_t)w * WSEG + (size_t)lane * V;

    VT x[kSegQ][V];
    KT kk[kSegQ][V + 1];  // kk[q][0] = key before the chunk
    uint32_t headm[kSegQ];
#pragma unroll
    for (int q = 0; q < kSegQ; ++q) {
        size_t p0 = wbase + (size_t)q * 64 * V;
        kk[q][0] = p0 > 0 && p0 - 1 < n ? keys[p0 - 1] : KT(0);
#pragma unroll
        for (int k = 0; k < V; ++k) {
            bool ok = p0 + k < n;
            x[q][k] = ok ? vals[p0 + k] : VT(0);
            kk[q][k + 1] = ok ? keys[p0 + k] : KT(0);
        }
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < V; ++k)
            if (p0 + k == 0 || !key_eq(kk[q][k], kk[q][k + 1])) m |= 1u << k;
        headm[q] = m;
    }
    // chunk folds of (head, value) pairs
    S tot[kSegQ];
#pragma unroll
    for (int q = 0; q < kSegQ; ++q) {
        size_t p0 = wbase + (size_t)q * 64 * V;
        S s{P::id(), 0u};
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (p0 + k < n) {
                if (headm[q] >> k & 1u) s = S{P::add(P::id(), x[q][k]), s.n + 1};
                else s.v = P::add(s.v, x[q][k]);
            }
        }
        tot[q] = s;
    }
    const S sid{P::id(), 0u};
    auto mrg = [](S a, S b) { return seg_merge<A, P>(a, b); };
    S carry = sid;
    S pre[kSegQ];
#pragma unroll
    for (int q = 0; q < kSegQ; ++q) {
        S inc = wave_inclusive_scan(tot[q], sid, mrg);
        S exc = wave_shr1(inc, sid);
        pre[q] = mrg(carry, exc);
        carry = mrg(carry, readlane(inc, 63));
    }
    if (lane == 0) s_wave[w] = carry;
    __syncthreads();
    S wpre = sid;
    for (uint32_t i = 0; i < w; ++i) wpre = mrg(wpre, s_wave[i]);
    if (w == 0) {
        S total = mrg(mrg(s_wave[0], s_wave[1]), mrg(s_wave[2], s_wave[3]));
        C agg{P::carry(total.v), total.n};
        C cid{P::cid(), 0u};
        C prefix = lookback_scan<C>(status, tile, epoch, agg, cid, cid, [](C a, C b) {
            return C{b.n ? b.v : P::cmerge(a.v, b.v), a.n + b.n};
        }, ctrl);
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    const C prefix = s_prefix;
#pragma unroll
    for (int q = 0; q < kSegQ; ++q) {
        size_t p0 = wbase + (size_t)q * 64 * V;
        S s = mrg(wpre, pre[q]);  // exclusive local pair before the chunk
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const size_t i = p0 + k;
            if (i >= n) break;
            const bool head = headm[q] >> k & 1u;
            if constexpr (MODE == SEG_RBK) {
                if (head && i > 0) {
                    // finished run i-1: exclusive pair at i
                    uint32_t nb = prefix.n + s.n;
                    if (nb >= 1 && nb <= n) {
                        keys_out[nb - 1] = kk[q][k];
                        out[nb - 1] = P::fin(prefix.v, s.v, s.n == 0);
                    } else {
                        __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            } else if constexpr (MODE == SEG_EXCL) {
                out[i] = head ? init : P::fin_init(init, prefix.v, s.v, s.n == 0);
            }
            if (head) s = S{P::add(P::id(), x[q][k]), s.n + 1};
            else s.v = P::add(s.v, x[q][k]);
            if constexpr (MODE == SEG_INCL) out[i] = P::fin(prefix.v, s.v, s.n == 0);
            if constexpr (MODE == SEG_RBK) {
                if (i == n - 1) {
                    uint32_t nb = prefix.n + s.n;
                    if (nb >= 1 && nb <= n) {
                        keys_out[nb - 1] = kk[q][k + 1];
                        out[nb - 1] = P::fin(prefix.v, s.v, s.n == 0);
                    } else {
                        __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    *count_out = nb;
                }
            }
        }
    }
}

template <typename KT, typename VT, typename P, int MODE, int Q>
static int launch_seg(bcmi_queue q, const void *keys, const void *vals, void *out, void *keys_out, size_t n,
                      uint64_t *status, VT ini, uint64_t *count_out)
{
    constexpr size_t TILE = (size_t)4 * Q * 64 * (16 / sizeof(VT));
    size_t tiles = (n + TILE - 1) / TILE;
    BCMI_TRY(begin_lookback(q, 1));
    BCMI_LAUNCH(q, MODE == SEG_RBK ? "reduce_by_key" : "segmented_scan",
                hipLaunchKernelGGL((k_segscan<KT, VT, P, MODE, Q>), dim3((unsigned)tiles), dim3(kSegThreads), 0,
                                   q->stream, (const KT *)keys, (const VT *)vals, (VT *)out, (KT *)keys_out, n,
                                   q->ctrl, status, ini, count_out, 1u));
    return 0;
}

// Chunks per thread from the tuning cache (key "__bcmi_segscan", param
// "q": 2, 4 or 8) for int32 keys with 4-byte values; 4 otherwise.
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
        const unsigned qq = tune_get("__bcmi_segscan", "q", 8);
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
        if (op == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED)
            return run_seg<KT, VT, SegCompF32, MODE>(q, keys, vals, out, keys_out, n, init, count_out);
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
