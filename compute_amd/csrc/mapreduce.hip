// mapreduce.hip -- transform_reduce / inner_product.
//
// Reference: transform_reduce (algorithm/transform_reduce.hpp) reduces a
// transform_iterator (over a zip_iterator for two ranges); inner_product
// (algorithm/inner_product.hpp) accumulates the products of a zipped pair,
// and its general form writes transform(first1, first2) to a temporary
// vector and accumulates that.  Float plus in accumulate is one work-item
// (accumulate.hpp:62-69).
//
// Here: one streaming pass of a few blocks per CU (as reduce.hip) evaluates
// the element map while folding, then one block folds the block partials
// in a fixed order.  No temporary.  Compensated float sums keep the exact
// rounding error of a final product or square (fma), so a float inner
// product is accurate to about the FP64 result.  Counting sequences are
// generated, not loaded.
#include "reduce.hpp"

namespace bcmi {

constexpr int kMapUnary = -1;  // OPB of a unary map (BCMI_PLUS is 0)

// The element map as the kernels see it.
struct MapDev {
    uint32_t binary;  // y = g(a op b); else y = g(a - c)
    uint32_t op;      // BCMI_PLUS, BCMI_MINUS, BCMI_MULTIPLIES
    uint32_t fn;      // BCMI_MAP_ID, _ABS, _SQUARE
    uint64_t c;
};

// Fold y = map(a, b) into the state.  FN is the map's g, OPB the binary op
// (kMapUnary: y = g(a - c)).  The last multiplication of a product or square goes
// through add_prod, so compensated policies keep its rounding error.
template <typename T, typename P, int FN, int OPB>
BDEV typename P::State map_fold(typename P::State s, T a, T b, T c)
{
    if constexpr (OPB == BCMI_MULTIPLIES) {
        if constexpr (FN == BCMI_MAP_ID) return P::add_prod(s, a, b);
        else if constexpr (FN == BCMI_MAP_ABS) return P::add_prod(s, map_abs(a), map_abs(b));
        else {
            const T t = apply_op<T, BCMI_MULTIPLIES>(a, b);
            return P::add_prod(s, t, t);
        }
    } else {
        T t;
        if constexpr (OPB == BCMI_PLUS) t = apply_op<T, BCMI_PLUS>(a, b);
        else if constexpr (OPB == BCMI_MINUS) t = map_sub(a, b);
        else t = map_sub(a, c);  // kMapUnary: a - c
        if constexpr (FN == BCMI_MAP_ID) return P::add(s, t);
        else if constexpr (FN == BCMI_MAP_ABS) return P::add(s, map_abs(t));
        else return P::add_prod(s, t, t);
    }
}

// element i of a counting sequence: start + i in the element type
template <typename T>
BDEV T count_at(int64_t start, size_t i)
{
    if constexpr (std::is_floating_point<T>::value) return (T)start + (T)i;
    else return (T)((uint64_t)start + (uint64_t)i);
}

// Two arrays (or one), both at the same offset inside a 16-byte chunk:
// block 0 takes the head and tail elements, all blocks stream the chunks,
// two in flight per thread.
template <typename T, typename P, int FN, int OPB>
BDEV typename P::State fold_arrays(const T *a, const T *b, size_t n, size_t head, size_t nchunks, T c)
{
    using S = typename P::State;
    constexpr int V = chunk16<T>::N;
    S acc = P::id();
    const size_t tail0 = head + nchunks * V;
    if (blockIdx.x == 0) {
        for (size_t i = threadIdx.x; i < head; i += kRedThreads)
            acc = map_fold<T, P, FN, OPB>(acc, a[i], OPB != kMapUnary ? b[i] : T(0), c);
        for (size_t i = tail0 + threadIdx.x; i < n; i += kRedThreads)
            acc = map_fold<T, P, FN, OPB>(acc, a[i], OPB != kMapUnary ? b[i] : T(0), c);
    }
    const chunk16<T> *ca = (const chunk16<T> *)(a + head);
    const chunk16<T> *cb = (const chunk16<T> *)(OPB != kMapUnary ? b + head : a + head);
    const size_t stride = (size_t)gridDim.x * kRedThreads;
    size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x;
    for (; i + stride < nchunks; i += 2 * stride) {
        const chunk16<T> x0 = ca[i], x1 = ca[i + stride];
        chunk16<T> y0{}, y1{};
        if constexpr (OPB != kMapUnary) {
            y0 = cb[i];
            y1 = cb[i + stride];
        }
#pragma unroll
        for (int k = 0; k < V; ++k) acc = map_fold<T, P, FN, OPB>(acc, x0.v[k], y0.v[k], c);
#pragma unroll
        for (int k = 0; k < V; ++k) acc = map_fold<T, P, FN, OPB>(acc, x1.v[k], y1.v[k], c);
    }
    for (; i < nchunks; i += stride) {
        const chunk16<T> x0 = ca[i];
        chunk16<T> y0{};
        if constexpr (OPB != kMapUnary) y0 = cb[i];
#pragma unroll
        for (int k = 0; k < V; ++k) acc = map_fold<T, P, FN, OPB>(acc, x0.v[k], y0.v[k], c);
    }
    return acc;
}

// Any mix of arrays and counting sequences, element by element.
template <typename T, typename P, int FN, int OPB>
BDEV typename P::State fold_general(const T *a, int64_t a0, const T *b, int64_t b0, size_t n, T c)
{
    using S = typename P::State;
    S acc = P::id();
    const size_t stride = (size_t)gridDim.x * kRedThreads;
    for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += stride) {
        const T x = a ? a[i] : count_at<T>(a0, i);
        const T y = OPB != kMapUnary ? (b ? b[i] : count_at<T>(b0, i)) : T(0);
        acc = map_fold<T, P, FN, OPB>(acc, x, y, c);
    }
    return acc;
}

template <typename T, typename P, bool VEC, int OPB, int FN>
BDEV typename P::State fold_any(const T *a, int64_t a0, const T *b, int64_t b0, size_t n, size_t head,
                                size_t nchunks, T c)
{
    if constexpr (VEC) return fold_arrays<T, P, FN, OPB>(a, b, n, head, nchunks, c);
    else return fold_general<T, P, FN, OPB>(a, a0, b, b0, n, c);
}

// VEC: both inputs are arrays sharing their offset inside a 16-byte chunk.
template <typename T, typename P, bool VEC, int OPB>
__global__ __launch_bounds__(kRedThreads) void k_mapreduce(const T *a, int64_t a0, const T *b, int64_t b0, size_t n,
                                                            size_t head, size_t nchunks, MapDev m,
                                                            typename P::Part *__restrict__ partials)
{
    using S = typename P::State;
    __shared__ S s_buf[kRedThreads];
    const T c = map_bits<T>(m.c);
    S acc;
    switch (m.fn) {
    case BCMI_MAP_ABS: acc = fold_any<T, P, VEC, OPB, BCMI_MAP_ABS>(a, a0, b, b0, n, head, nchunks, c); break;
    case BCMI_MAP_SQUARE: acc = fold_any<T, P, VEC, OPB, BCMI_MAP_SQUARE>(a, a0, b, b0, n, head, nchunks, c); break;
    default: acc = fold_any<T, P, VEC, OPB, BCMI_MAP_ID>(a, a0, b, b0, n, head, nchunks, c); break;
    }
    S blk = block_combine<S, P>(acc, s_buf);
    if (threadIdx.x == 0) partials[blockIdx.x] = P::part(blk);
}

template <typename T, typename P, int OPB>
static int launch_mapreduce(bcmi_queue q, const bcmi_seq *a, const bcmi_seq *b, size_t n, const MapDev &m,
                            const void *init, void *out, double *out_f64)
{
    const T *pa = (const T *)a->ptr;
    constexpr bool BIN = OPB != kMapUnary;
    const T *pb = BIN ? (const T *)b->ptr : nullptr;
    // the chunked path needs arrays at one offset inside a 16-byte chunk;
    // 8- and 16-bit elements always go element by element (the compiler's
    // vectoriser fails on the 16-wide byte chunks of some maps)
    const uintptr_t ma = (uintptr_t)pa & 15, mb = BIN ? ((uintptr_t)pb & 15) : ma;
    const bool vec = sizeof(T) >= 4 && pa && (!BIN || pb) && ma == mb && ma % sizeof(T) == 0;
    size_t head = 0, nchunks = 0;
    if (vec) {
        head = ma ? (16 - ma) / sizeof(T) : 0;
        if (head > n) head = n;
        nchunks = (n - head) / chunk16<T>::N;
    }
    const size_t work = vec ? nchunks : (n + 3) / 4;
    const size_t want = (work + kRedThreads * 8 - 1) / (kRedThreads * 8);
    const size_t maxb = (size_t)q->cu_count * 4;
    const size_t blocks = want < 1 ? 1 : (want > maxb ? maxb : want);
    void *scr;
    BCMI_TRY(scratch(q, blocks * sizeof(typename P::Part), &scr));
    T ini{};
    if (init) memcpy(&ini, init, sizeof(T));
    const int64_t a0 = a->start, b0 = BIN ? b->start : 0;
    if constexpr (sizeof(T) >= 4) {
        if (vec) {
            BCMI_LAUNCH(q, "mapreduce",
                        hipLaunchKernelGGL((k_mapreduce<T, P, true, OPB>), dim3((unsigned)blocks), dim3(kRedThreads),
                                           0, q->stream, pa, a0, pb, b0, n, head, nchunks, m,
                                           (typename P::Part *)scr));
        }
    }
    if (!vec) {
        BCMI_LAUNCH(q, "mapreduce",
                    hipLaunchKernelGGL((k_mapreduce<T, P, false, OPB>), dim3((unsigned)blocks), dim3(kRedThreads), 0,
                                       q->stream, pa, a0, pb, b0, n, head, nchunks, m,
                                       (typename P::Part *)scr));
    }
    BCMI_LAUNCH(q, "mapreduce_final",
                hipLaunchKernelGGL((k_reduce_final<T, P>), dim3(1), dim3(kRedThreads), 0, q->stream,
                                   (const typename P::Part *)scr, (uint32_t)blocks, ini, init ? 1 : 0, (T *)out,
                                   out_f64));
    return 0;
}

template <typename T, typename P>
static int mapreduce_op(bcmi_queue q, const bcmi_seq *a, const bcmi_seq *b, size_t n, const MapDev &m,
                        const void *init, void *out, double *o64)
{
    if (!m.binary) return launch_mapreduce<T, P, kMapUnary>(q, a, b, n, m, init, out, o64);
    switch (m.op) {
    case BCMI_PLUS: return launch_mapreduce<T, P, BCMI_PLUS>(q, a, b, n, m, init, out, o64);
    case BCMI_MINUS: return launch_mapreduce<T, P, BCMI_MINUS>(q, a, b, n, m, init, out, o64);
    default: return launch_mapreduce<T, P, BCMI_MULTIPLIES>(q, a, b, n, m, init, out, o64);
    }
}

template <typename T>
static int mapreduce_type(bcmi_queue q, int rop, const bcmi_seq *a, const bcmi_seq *b, size_t n, const MapDev &m,
                          const void *init, void *out, double *o64, int precision)
{
    if constexpr (std::is_same<T, float>::value) {
        if (rop == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED)
            return mapreduce_op<float, CompF32Policy>(q, a, b, n, m, init, out, o64);
    }
    if constexpr (std::is_same<T, double>::value) {
        if (rop == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED)
            return mapreduce_op<double, CompF64Policy>(q, a, b, n, m, init, out, o64);
    }
    switch (rop) {
    case BCMI_PLUS: return mapreduce_op<T, PlainPolicy<T, BCMI_PLUS>>(q, a, b, n, m, init, out, o64);
    case BCMI_MULTIPLIES: return mapreduce_op<T, PlainPolicy<T, BCMI_MULTIPLIES>>(q, a, b, n, m, init, out, o64);
    case BCMI_MIN: return mapreduce_op<T, PlainPolicy<T, BCMI_MIN>>(q, a, b, n, m, init, out, o64);
    case BCMI_MAX: return mapreduce_op<T, PlainPolicy<T, BCMI_MAX>>(q, a, b, n, m, init, out, o64);
    default: set_error("transform_reduce: unsupported reduce op %d", rop); return BCMI_INVALID_VALUE;
    }
}

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_transform_reduce(bcmi_queue q, int dtype, const bcmi_seq *a, const bcmi_seq *b, size_t n,
                                     const bcmi_map *map, int reduce_op, const void *init, void *out,
                                     double *out_f64, int precision)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(a && map && out, BCMI_INVALID_VALUE, "transform_reduce: null argument");
    BCMI_CHECK(dtype_valid(dtype), BCMI_INVALID_VALUE, "transform_reduce: bad dtype %d", dtype);
    BCMI_CHECK(map->fn >= BCMI_MAP_ID && map->fn <= BCMI_MAP_SQUARE, BCMI_INVALID_VALUE,
               "transform_reduce: bad map fn %d", map->fn);
    BCMI_CHECK(!b || map->op == BCMI_PLUS || map->op == BCMI_MINUS || map->op == BCMI_MULTIPLIES,
               BCMI_INVALID_VALUE, "transform_reduce: bad map op %d", map->op);
    if (n == 0) {
        // an empty range leaves the result untouched (reduce.hpp:282-284);
        // with an init, accumulate's result is init (accumulate.hpp:45-48)
        if (init) BCMI_TRY(bcmi_memcpy(q, out, init, (size_t)dtype_size(dtype), BCMI_H2D, 1));
        return 0;
    }
    MapDev m;
    m.binary = b ? 1u : 0u;
    m.op = (uint32_t)map->op;
    m.fn = (uint32_t)map->fn;
    m.c = b ? 0 : map->offset;
    switch (dtype) {
    case BCMI_I8: return mapreduce_type<int8_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_U8: return mapreduce_type<uint8_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_I16: return mapreduce_type<int16_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_U16: return mapreduce_type<uint16_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_I32: return mapreduce_type<int32_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_U32: return mapreduce_type<uint32_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_I64: return mapreduce_type<int64_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_U64: return mapreduce_type<uint64_t>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    case BCMI_F32: return mapreduce_type<float>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    default: return mapreduce_type<double>(q, reduce_op, a, b, n, m, init, out, out_f64, precision);
    }
}
