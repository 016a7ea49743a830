// reduce.hip -- reduce / accumulate.
//
// Reference: reduce_on_gpu (algorithm/detail/reduce_on_gpu.hpp:169-280):
// VPT=8 values per work-item strided by TPB=128, an LDS pairwise tree per
// group, then extra passes until one value remains (several launches).
// accumulate(float, plus) is a one-work-item left fold
// (accumulate.hpp:62-69, serial_accumulate.hpp:38-40).
//
// Here: one launch.  A grid of a few blocks per CU streams the input with
// 128-bit loads; each thread keeps a compensated FP32 pair (TwoSum) for
// float plus, waves combine pairs with DPP, the block writes one FP64
// partial, and a one-block kernel folds the partials in FP64 in a fixed
// order (so the result is deterministic).  Integers wrap exactly;
// min/max/multiplies fold in the element type.
#include "device.hpp"
#include "ops.hpp"

#include <cstring>
#include <limits>
#include <type_traits>

namespace bcmi {

constexpr int kRedThreads = 256;

// Exact policy: fold in the element type (integers; min/max; and the naive
// float tree used only for error comparison).
template <typename T, int OP>
struct PlainPolicy {
    using State = T;
    using Part = T;
    static BDEV State id() { return op_identity<T, OP>(); }
    static BDEV State add(State s, T x) { return apply_op<T, OP>(s, x); }
    static BDEV State merge(State a, State b) { return apply_op<T, OP>(a, b); }
    static BDEV Part part(State s) { return s; }
    static BDEV Part part_id() { return id(); }
    static BDEV Part part_merge(Part a, Part b) { return merge(a, b); }
    static BDEV T finish(Part p, bool has_init, T init) { return has_init ? apply_op<T, OP>(init, p) : p; }
    static BDEV double f64(Part p) { return (double)p; }
};

// Compensated float plus: TwoSum pairs in FP32, FP64 partials.
struct CompF32Policy {
    using State = cf32;
    using Part = double;
    static BDEV State id() { return cf32{0.f, 0.f}; }
    static BDEV State add(State s, float x) { return cf_add(s, x); }
    static BDEV State merge(State a, State b) { return cf_merge(a, b); }
    static BDEV Part part(State s) { return cf_value(s); }
    static BDEV Part part_id() { return 0.0; }
    static BDEV Part part_merge(Part a, Part b) { return a + b; }
    static BDEV float finish(Part p, bool has_init, float init) { return (float)(has_init ? (double)init + p : p); }
    static BDEV double f64(Part p) { return p; }
};

// Compensated double plus: TwoSum in FP64.
struct cf64 {
    double s, c;
};
struct CompF64Policy {
    using State = cf64;
    using Part = cf64;
    static BDEV void ts(double a, double b, double &s, double &e)
    {
        s = a + b;
        double bp = s - a;
        e = (a - (s - bp)) + (b - bp);
    }
    static BDEV State id() { return cf64{0.0, 0.0}; }
    static BDEV State add(State a, double x)
    {
        double s, e;
        ts(a.s, x, s, e);
        return cf64{s, a.c + e};
    }
    static BDEV State merge(State a, State b)
    {
        double s, e;
        ts(a.s, b.s, s, e);
        return cf64{s, (a.c + b.c) + e};
    }
    static BDEV Part part(State s) { return s; }
    static BDEV Part part_id() { return id(); }
    static BDEV Part part_merge(Part a, Part b) { return merge(a, b); }
    static BDEV double finish(Part p, bool has_init, double init)
    {
        if (has_init) p = add(p, init);
        return p.s + p.c;
    }
    static BDEV double f64(Part p) { return p.s + p.c; }
};

// 16-byte-chunk view of the input
template <typename T>
struct alignas(16) chunk16 {
    static constexpr int N = 16 / sizeof(T);
    T v[N];
};

// DPP moves 4- or 8-byte values; 16-byte states (cf64) go through LDS.
template <typename S, typename P>
BDEV S block_combine(S v, S *s_buf)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if constexpr (sizeof(S) == 4 || sizeof(S) == 8) {
        v = wave_reduce(v, P::id(), [](S a, S b) { return P::merge(a, b); });
        if (lane == 0) s_buf[w] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            S acc = s_buf[0];
            for (int i = 1; i < kRedThreads / 64; ++i) acc = P::merge(acc, s_buf[i]);
            s_buf[0] = acc;
        }
        __syncthreads();
        return s_buf[0];
    } else {
        s_buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = kRedThreads / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) s_buf[threadIdx.x] = P::merge(s_buf[threadIdx.x], s_buf[threadIdx.x + off]);
            __syncthreads();
        }
        return s_buf[0];
    }
}

// U chunks of 16 bytes in flight per thread and iteration.
template <typename T, typename P, int U>
__global__ __launch_bounds__(kRedThreads) void k_reduce(const T *__restrict__ in, size_t n, size_t head,
                                                        size_t nchunks, typename P::Part *__restrict__ partials)
{
    using S = typename P::State;
    __shared__ S s_buf[kRedThreads];

    S acc = P::id();
    // head/tail elements outside the aligned 16-byte body: block 0
    const size_t tail0 = head + nchunks * chunk16<T>::N;
    if (blockIdx.x == 0) {
        for (size_t i = threadIdx.x; i < head; i += kRedThreads) acc = P::add(acc, in[i]);
        for (size_t i = tail0 + threadIdx.x; i < n; i += kRedThreads) acc = P::add(acc, in[i]);
    }
    const chunk16<T> *body = (const chunk16<T> *)(in + head);
    const size_t stride = (size_t)gridDim.x * kRedThreads;
    size_t c = (size_t)blockIdx.x * kRedThreads + threadIdx.x;
    for (; c + (U - 1) * stride < nchunks; c += U * stride) {
        chunk16<T> a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = body[c + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < chunk16<T>::N; ++k) acc = P::add(acc, a[u].v[k]);
    }
    for (; c < nchunks; c += stride) {
        const chunk16<T> a = body[c];
#pragma unroll
        for (int k = 0; k < chunk16<T>::N; ++k) acc = P::add(acc, a.v[k]);
    }
    S blk = block_combine<S, P>(acc, s_buf);
    if (threadIdx.x == 0) partials[blockIdx.x] = P::part(blk);
}

// One block folds the per-block partials in FP64 / the element type in a
// fixed order, so the result is deterministic.
template <typename T, typename P>
__global__ __launch_bounds__(kRedThreads) void k_reduce_final(const typename P::Part *__restrict__ partials,
                                                              uint32_t nparts, T init, int has_init, T *out,
                                                              double *out_f64)
{
    using Pt = typename P::Part;
    __shared__ Pt s_part[kRedThreads];
    Pt p = P::part_id();
    for (uint32_t i = threadIdx.x; i < nparts; i += kRedThreads) p = P::part_merge(p, partials[i]);
    s_part[threadIdx.x] = p;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_part[threadIdx.x] = P::part_merge(s_part[threadIdx.x], s_part[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Pt total = s_part[0];
        *out = P::finish(total, has_init != 0, init);
        if (out_f64) *out_f64 = P::f64(total) + (has_init ? (double)init : 0.0);
    }
}

template <typename T, typename P>
static int run_reduce(bcmi_queue q, const void *in_, size_t n, const void *init, void *out, double *out_f64)
{
    const T *in = (const T *)in_;
    size_t head = 0;
    uintptr_t mis = (uintptr_t)in & 15;
    if (mis) {
        if (mis % sizeof(T)) head = n;  // not element-aligned to a chunk: scalar only
        else head = (16 - mis) / sizeof(T);
    }
    if (head > n) head = n;
    size_t nchunks = (n - head) / chunk16<T>::N;
    // chunks in flight per thread ("unroll": 2, 4 or 8) and workgroups per
    // CU ("blocks_per_cu") come from the tuning cache (key "__bcmi_reduce")
    const unsigned unroll = tune_get("__bcmi_reduce", "unroll", 2);
    const unsigned bpc = tune_get("__bcmi_reduce", "blocks_per_cu", 4);
    size_t want = (nchunks + kRedThreads * 8 - 1) / (kRedThreads * 8);
    size_t maxb = (size_t)q->cu_count * (bpc ? bpc : 4);
    size_t blocks = want < 1 ? 1 : (want > maxb ? maxb : want);
    void *scr;
    BCMI_TRY(scratch(q, blocks * sizeof(typename P::Part), &scr));
    T ini{};
    if (init) memcpy(&ini, init, sizeof(T));
    auto go = [&](auto u) {
        BCMI_LAUNCH(q, "reduce",
                    hipLaunchKernelGGL((k_reduce<T, P, decltype(u)::value>), dim3((unsigned)blocks), dim3(kRedThreads),
                                       0, q->stream, in, n, head, nchunks, (typename P::Part *)scr));
        return 0;
    };
    if (unroll >= 8) BCMI_TRY(go(std::integral_constant<int, 8>()));
    else if (unroll >= 4) BCMI_TRY(go(std::integral_constant<int, 4>()));
    else BCMI_TRY(go(std::integral_constant<int, 2>()));
    BCMI_LAUNCH(q, "reduce_final",
                hipLaunchKernelGGL((k_reduce_final<T, P>), dim3(1), dim3(kRedThreads), 0, q->stream,
                                   (const typename P::Part *)scr, (uint32_t)blocks, ini, init ? 1 : 0, (T *)out,
                                   out_f64));
    return 0;
}

template <typename T>
static int reduce_int(bcmi_queue q, int op, const void *in, size_t n, const void *init, void *out, double *o64)
{
    switch (op) {
    case BCMI_PLUS: return run_reduce<T, PlainPolicy<T, BCMI_PLUS>>(q, in, n, init, out, o64);
    case BCMI_MULTIPLIES: return run_reduce<T, PlainPolicy<T, BCMI_MULTIPLIES>>(q, in, n, init, out, o64);
    case BCMI_MIN: return run_reduce<T, PlainPolicy<T, BCMI_MIN>>(q, in, n, init, out, o64);
    case BCMI_MAX: return run_reduce<T, PlainPolicy<T, BCMI_MAX>>(q, in, n, init, out, o64);
    default: set_error("reduce: unsupported op %d", op); return BCMI_INVALID_VALUE;
    }
}

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_reduce(bcmi_queue q, int dtype, int op, const void *in, size_t n, const void *init, void *out,
                           double *out_f64, int precision)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(out, BCMI_INVALID_VALUE, "reduce: null output");
    BCMI_CHECK(n == 0 || in, BCMI_INVALID_VALUE, "reduce: null input");
    if (n == 0) {
        // reduce() of an empty range leaves the result untouched
        // (reduce.hpp:282-284); accumulate() returns init (accumulate.hpp:45-48).
        if (init) BCMI_TRY(bcmi_memcpy(q, out, init, (size_t)dtype_size(dtype), BCMI_H2D, 1));
        return 0;
    }
    switch (dtype) {
    case BCMI_I8: return reduce_int<int8_t>(q, op, in, n, init, out, out_f64);
    case BCMI_U8: return reduce_int<uint8_t>(q, op, in, n, init, out, out_f64);
    case BCMI_I16: return reduce_int<int16_t>(q, op, in, n, init, out, out_f64);
    case BCMI_U16: return reduce_int<uint16_t>(q, op, in, n, init, out, out_f64);
    case BCMI_I32: return reduce_int<int32_t>(q, op, in, n, init, out, out_f64);
    case BCMI_U32: return reduce_int<uint32_t>(q, op, in, n, init, out, out_f64);
    case BCMI_I64: return reduce_int<int64_t>(q, op, in, n, init, out, out_f64);
    case BCMI_U64: return reduce_int<uint64_t>(q, op, in, n, init, out, out_f64);
    case BCMI_F32:
        if (op == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED)
            return run_reduce<float, CompF32Policy>(q, in, n, init, out, out_f64);
        return reduce_int<float>(q, op, in, n, init, out, out_f64);
    case BCMI_F64:
        if (op == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED)
            return run_reduce<double, CompF64Policy>(q, in, n, init, out, out_f64);
        return reduce_int<double>(q, op, in, n, init, out, out_f64);
    default: set_error("reduce: bad dtype %d", dtype); return BCMI_INVALID_VALUE;
    }
}
