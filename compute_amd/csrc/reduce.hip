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
#include "reduce.hpp"

#include <cstring>
#include <limits>
#include <type_traits>

namespace bcmi {

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
