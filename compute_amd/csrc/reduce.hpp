// reduce.hpp -- the reduction policies and the block / final combine shared
// by reduce.hip (reduce, accumulate) and mapreduce.hip (transform_reduce,
// inner_product).
//
// A policy P says how one thread folds elements (State), what a block
// writes as its partial (Part, FP64 for compensated float sums) and how the
// one-block final kernel folds the partials in a fixed order, so results
// are deterministic.  add_prod(s, x, y) folds the product x * y; the
// compensated policies keep that product's rounding error exactly (fma).
#pragma once

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
    static BDEV State add_prod(State s, T x, T y) { return add(s, apply_op<T, BCMI_MULTIPLIES>(x, y)); }
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
    static BDEV State add_prod(State s, float x, float y)
    {
        float p = x * y;
        // p must stay the rounded product: fused into TwoSum's additions
        // (contraction) it would carry the rounding error a second time
        asm volatile("" : "+v"(p));
        State r = cf_add(s, p);
        r.c += __builtin_fmaf(x, y, -p);  // the product's exact rounding error
        return r;
    }
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
    static BDEV State add_prod(State s, double x, double y)
    {
        double p = x * y;
        asm volatile("" : "+v"(p));  // the rounded product, not contracted (see CompF32Policy)
        State r = add(s, p);
        r.c += __builtin_fma(x, y, -p);
        return r;
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
        block_barrier();
        if (threadIdx.x == 0) {
            S acc = s_buf[0];
            for (int i = 1; i < kRedThreads / 64; ++i) acc = P::merge(acc, s_buf[i]);
            s_buf[0] = acc;
        }
        block_barrier();
        return s_buf[0];
    } else {
        s_buf[threadIdx.x] = v;
        block_barrier();
        for (int off = kRedThreads / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) s_buf[threadIdx.x] = P::merge(s_buf[threadIdx.x], s_buf[threadIdx.x + off]);
            block_barrier();
        }
        return s_buf[0];
    }
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
    block_barrier();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_part[threadIdx.x] = P::part_merge(s_part[threadIdx.x], s_part[threadIdx.x + off]);
        block_barrier();
    }
    if (threadIdx.x == 0) {
        Pt total = s_part[0];
        *out = P::finish(total, has_init != 0, init);
        if (out_f64) *out_f64 = P::f64(total) + (has_init ? (double)init : 0.0);
    }
}

}  // namespace bcmi
