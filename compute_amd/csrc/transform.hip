// transform.hip -- elementwise transform (saxpy and binary functors).
//
// Reference: transform.hpp:38-82 -> copy(transform_iterator) ->
// copy_on_device_gpu (detail/copy_on_device.hpp:59-92), VPT=4 strided by
// TPB=128.  Here: 128-bit loads/stores, each thread moving U 16-byte chunks
// of every operand, a grid that covers the array once (HBM-bound:
// saxpy moves 12 B per element).
#include "device.hpp"
#include "ops.hpp"

#include <type_traits>

namespace bcmi {

constexpr int kTfThreads = 256;
constexpr int kTfUnroll = 4;

template <typename T>
struct vec16 {
    static constexpr int N = 16 / sizeof(T);
    T v[N];
};

struct SaxpyOp {
    float a;
    BDEV float operator()(float x, float y) const { return __builtin_fmaf(a, x, y); }
};

template <typename T, bool = std::is_integral<T>::value>
struct wrap_type { using type = T; };
template <typename T>
struct wrap_type<T, true> { using type = std::make_unsigned_t<T>; };

template <typename T, int OP>
struct BinOp {
    // integers wrap modulo 2^W (OpenCL C semantics): do +, -, * unsigned
    using U = typename wrap_type<T>::type;
    BDEV T operator()(T x, T y) const
    {
        if constexpr (OP == BCMI_PLUS) return (T)(U)((U)x + (U)y);
        else if constexpr (OP == BCMI_MINUS) return (T)(U)((U)x - (U)y);
        else if constexpr (OP == BCMI_MULTIPLIES) return (T)(U)((U)x * (U)y);
        else if constexpr (OP == BCMI_DIVIDES) return (T)(x / y);
        else if constexpr (OP == BCMI_MIN) return y < x ? y : x;
        else return x < y ? y : x;
    }
};

// Vector body over whole 16-byte chunks; the caller guarantees a, b and out
// are 16-byte aligned.
template <typename T, typename F>
__global__ __launch_bounds__(kTfThreads) void k_transform_vec(const vec16<T> *__restrict__ a,
                                                              const vec16<T> *__restrict__ b,
                                                              vec16<T> *__restrict__ out, size_t nchunks, F f)
{
    size_t base = (size_t)blockIdx.x * (kTfThreads * kTfUnroll) + threadIdx.x;
    vec16<T> va[kTfUnroll], vb[kTfUnroll];
#pragma unroll
    for (int j = 0; j < kTfUnroll; ++j) {
        size_t c = base + (size_t)j * kTfThreads;
        if (c < nchunks) {
            va[j] = a[c];
            vb[j] = b[c];
        }
    }
#pragma unroll
    for (int j = 0; j < kTfUnroll; ++j) {
        size_t c = base + (size_t)j * kTfThreads;
        if (c < nchunks) {
            vec16<T> r;
#pragma unroll
            for (int k = 0; k < vec16<T>::N; ++k) r.v[k] = f(va[j].v[k], vb[j].v[k]);
            out[c] = r;
        }
    }
}

// Scalar path for heads/tails and for operands that are not mutually aligned.
template <typename T, typename F>
__global__ __launch_bounds__(kTfThreads) void k_transform_scalar(const T *__restrict__ a, const T *__restrict__ b,
                                                                 T *__restrict__ out, size_t n, F f)
{
    size_t i = (size_t)blockIdx.x * kTfThreads + threadIdx.x;
    size_t stride = (size_t)gridDim.x * kTfThreads;
    for (; i < n; i += stride) out[i] = f(a[i], b[i]);
}

template <typename T, typename F>
static int run_transform(bcmi_queue q, const T *a, const T *b, T *out, size_t n, F f, const char *name)
{
    if (n == 0) return 0;
    const uintptr_t mask = 15;
    size_t head = 0;
    uintptr_t oa = (uintptr_t)out & mask;
    bool aligned = (((uintptr_t)a & mask) == oa) && (((uintptr_t)b & mask) == oa) && (oa % sizeof(T) == 0);
    if (aligned && oa != 0) head = (16 - oa) / sizeof(T);
    if (head > n) head = n;
    size_t body = aligned ? (n - head) / vec16<T>::N : 0;
    size_t done = head + body * vec16<T>::N;
    int maxb = q->cu_count * 8;
    if (!aligned) {
        size_t blocks = (n + kTfThreads - 1) / kTfThreads;
        if (blocks > (size_t)maxb) blocks = maxb;
        BCMI_LAUNCH(q, name, hipLaunchKernelGGL((k_transform_scalar<T, F>), dim3((unsigned)blocks), dim3(kTfThreads), 0,
                                                q->stream, a, b, out, n, f));
        return 0;
    }
    if (head) {
        BCMI_LAUNCH(q, name, hipLaunchKernelGGL((k_transform_scalar<T, F>), dim3(1), dim3(kTfThreads), 0, q->stream, a,
                                                b, out, head, f));
    }
    if (body) {
        size_t per = (size_t)kTfThreads * kTfUnroll;
        size_t blocks = (body + per - 1) / per;
        BCMI_LAUNCH(q, name,
                    hipLaunchKernelGGL((k_transform_vec<T, F>), dim3((unsigned)blocks), dim3(kTfThreads), 0, q->stream,
                                       (const vec16<T> *)(a + head), (const vec16<T> *)(b + head),
                                       (vec16<T> *)(out + head), body, f));
    }
    if (done < n) {
        BCMI_LAUNCH(q, name, hipLaunchKernelGGL((k_transform_scalar<T, F>), dim3(1), dim3(kTfThreads), 0, q->stream,
                                                a + done, b + done, out + done, n - done, f));
    }
    return 0;
}

template <typename T>
static int transform_typed(bcmi_queue q, int op, const void *a, const void *b, void *out, size_t n)
{
    const T *pa = (const T *)a, *pb = (const T *)b;
    T *po = (T *)out;
    switch (op) {
    case BCMI_PLUS: return run_transform(q, pa, pb, po, n, BinOp<T, BCMI_PLUS>(), "transform");
    case BCMI_MINUS: return run_transform(q, pa, pb, po, n, BinOp<T, BCMI_MINUS>(), "transform");
    case BCMI_MULTIPLIES: return run_transform(q, pa, pb, po, n, BinOp<T, BCMI_MULTIPLIES>(), "transform");
    case BCMI_DIVIDES: return run_transform(q, pa, pb, po, n, BinOp<T, BCMI_DIVIDES>(), "transform");
    case BCMI_MIN: return run_transform(q, pa, pb, po, n, BinOp<T, BCMI_MIN>(), "transform");
    case BCMI_MAX: return run_transform(q, pa, pb, po, n, BinOp<T, BCMI_MAX>(), "transform");
    default: set_error("transform: unsupported op %d", op); return BCMI_INVALID_VALUE;
    }
}

}  // namespace bcmi

using namespace bcmi;

// experimental::sort_by_transform's key extraction: one key_dtype field of
// each record, optionally |x| (integers wrap).
template <typename K>
__global__ __launch_bounds__(kTfThreads) void k_extract_keys(const unsigned char *__restrict__ rec, size_t rb,
                                                             size_t off, size_t n, int fn, K *__restrict__ keys)
{
    const size_t stride = (size_t)gridDim.x * kTfThreads;
    for (size_t i = (size_t)blockIdx.x * kTfThreads + threadIdx.x; i < n; i += stride) {
        K k = *(const K *)(rec + i * rb + off);
        if (fn == BCMI_MAP_ABS) {
            if constexpr (std::is_floating_point<K>::value) {
                k = __builtin_fabs(k);
            } else if constexpr (std::is_signed<K>::value) {
                using U = typename wrap_t<K>::type;
                if (k < 0) k = (K)(U)(0u - (U)k);
            }
        }
        keys[i] = k;
    }
}

template <typename K>
static int extract_keys(bcmi_queue q, int fn, const void *rec, size_t rb, size_t off, size_t n, void *keys)
{
    BCMI_CHECK(off + sizeof(K) <= rb, BCMI_INVALID_VALUE, "extract_keys: field past the record");
    BCMI_CHECK(((uintptr_t)rec + off) % sizeof(K) == 0 && rb % sizeof(K) == 0, BCMI_INVALID_VALUE,
               "extract_keys: key field not aligned to its type");
    const unsigned blocks = (unsigned)std::min<size_t>((n + kTfThreads - 1) / kTfThreads, (size_t)cu_blocks(q, 8));
    BCMI_LAUNCH(q, "extract_keys",
                hipLaunchKernelGGL((k_extract_keys<K>), dim3(blocks), dim3(kTfThreads), 0, q->stream,
                                   (const unsigned char *)rec, rb, off, n, fn, (K *)keys));
    return 0;
}

extern "C" int bcmi_extract_keys(bcmi_queue q, int key_dtype, int fn, const void *records, size_t record_bytes,
                                 size_t key_offset, size_t n, void *keys)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(fn == BCMI_MAP_ID || fn == BCMI_MAP_ABS, BCMI_INVALID_VALUE, "extract_keys: bad fn %d", fn);
    if (n == 0) return 0;
    BCMI_CHECK(records && keys, BCMI_INVALID_VALUE, "extract_keys: null pointer");
    switch (key_dtype) {
    case BCMI_I8: return extract_keys<int8_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_U8: return extract_keys<uint8_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_I16: return extract_keys<int16_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_U16: return extract_keys<uint16_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_I32: return extract_keys<int32_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_U32: return extract_keys<uint32_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_I64: return extract_keys<int64_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_U64: return extract_keys<uint64_t>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_F32: return extract_keys<float>(q, fn, records, record_bytes, key_offset, n, keys);
    case BCMI_F64: return extract_keys<double>(q, fn, records, record_bytes, key_offset, n, keys);
    default: set_error("extract_keys: bad dtype %d", key_dtype); return BCMI_INVALID_VALUE;
    }
}

extern "C" int bcmi_saxpy(bcmi_queue q, float a, const float *x, const float *y, float *out, size_t n)
{
    BCMI_TRY(use_queue(q));
    if (n == 0) return 0;
    BCMI_CHECK(x && y && out, BCMI_INVALID_VALUE, "saxpy: null pointer");
    return run_transform(q, x, y, out, n, SaxpyOp{a}, "saxpy");
}

extern "C" int bcmi_transform(bcmi_queue q, int dtype, int op, const void *a, const void *b, void *out, size_t n)
{
    BCMI_TRY(use_queue(q));
    if (n == 0) return 0;
    BCMI_CHECK(a && b && out, BCMI_INVALID_VALUE, "transform: null pointer");
    switch (dtype) {
    case BCMI_I8: return transform_typed<int8_t>(q, op, a, b, out, n);
    case BCMI_U8: return transform_typed<uint8_t>(q, op, a, b, out, n);
    case BCMI_I16: return transform_typed<int16_t>(q, op, a, b, out, n);
    case BCMI_U16: return transform_typed<uint16_t>(q, op, a, b, out, n);
    case BCMI_I32: return transform_typed<int32_t>(q, op, a, b, out, n);
    case BCMI_U32: return transform_typed<uint32_t>(q, op, a, b, out, n);
    case BCMI_I64: return transform_typed<int64_t>(q, op, a, b, out, n);
    case BCMI_U64: return transform_typed<uint64_t>(q, op, a, b, out, n);
    case BCMI_F32: return transform_typed<float>(q, op, a, b, out, n);
    case BCMI_F64: return transform_typed<double>(q, op, a, b, out, n);
    default: set_error("transform: bad dtype %d", dtype); return BCMI_INVALID_VALUE;
    }
}

// ------------------------------------------------ fancy-iterator inputs
// out[i] = op(a_i, b_i) (or a_i alone) where each operand is an array, a
// counting or a constant sequence, mapped element by element (IterIn).  A
// grid-stride loop: generated operands read nothing, arrays are read with
// coalesced scalar loads.
template <typename T, int OP>
__global__ __launch_bounds__(kTfThreads) void k_transform_iter(const IterIn<T> a, const IterIn<T> b, int binary,
                                                               T *__restrict__ out, size_t n)
{
    const size_t stride = (size_t)gridDim.x * kTfThreads;
    for (size_t i = (size_t)blockIdx.x * kTfThreads + threadIdx.x; i < n; i += stride) {
        const T x = a.at(i);
        out[i] = binary ? BinOp<T, OP>()(x, b.at(i)) : x;
    }
}

// length<floatW_>: the Euclidean length of each W-component float element
template <int W>
__global__ __launch_bounds__(kTfThreads) void k_vector_length(const float *__restrict__ a, float *__restrict__ out,
                                                              size_t n)
{
    const size_t stride = (size_t)gridDim.x * kTfThreads;
    for (size_t i = (size_t)blockIdx.x * kTfThreads + threadIdx.x; i < n; i += stride) {
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < W; ++c) s = __builtin_fmaf(a[i * W + c], a[i * W + c], s);
        out[i] = __builtin_sqrtf(s);
    }
}

template <typename T>
static int transform_iter_typed(bcmi_queue q, int op, const bcmi_iter *a, const bcmi_iter *b, void *out, size_t n)
{
    const IterIn<T> ia = iter_in<T>(a);
    IterIn<T> ib = ia;
    if (b) ib = iter_in<T>(b);
    const int binary = b ? 1 : 0;
    const unsigned blocks = (unsigned)std::min<size_t>((n + kTfThreads - 1) / kTfThreads, (size_t)cu_blocks(q, 8));
    T *o = (T *)out;
#define BCMI_TF_ITER(OPV)                                                                                       \
    BCMI_LAUNCH(q, "transform_iter", hipLaunchKernelGGL((k_transform_iter<T, OPV>), dim3(blocks), dim3(kTfThreads), \
                                                        0, q->stream, ia, ib, binary, o, n));                   \
    return 0
    switch (op) {
    case BCMI_PLUS: BCMI_TF_ITER(BCMI_PLUS);
    case BCMI_MINUS: BCMI_TF_ITER(BCMI_MINUS);
    case BCMI_MULTIPLIES: BCMI_TF_ITER(BCMI_MULTIPLIES);
    case BCMI_DIVIDES: BCMI_TF_ITER(BCMI_DIVIDES);
    case BCMI_MIN: BCMI_TF_ITER(BCMI_MIN);
    case BCMI_MAX: BCMI_TF_ITER(BCMI_MAX);
    default: set_error("transform: unsupported op %d", op); return BCMI_INVALID_VALUE;
    }
#undef BCMI_TF_ITER
}

// Element type conversion, OpenCL C's implicit conversion of the
// reference's copy and scan kernels (`result[i] = x` with differing element
// types, detail/copy_on_device.hpp:59-92, scan_on_gpu.hpp:86-88): integers
// wrap, int -> float rounds to nearest, float -> int truncates toward zero
// (out-of-range values are undefined there; here the hardware saturates).
template <typename S, typename D>
__global__ __launch_bounds__(kTfThreads) void k_convert(const S *__restrict__ src, D *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * kTfThreads;
    for (size_t i = (size_t)blockIdx.x * kTfThreads + threadIdx.x; i < n; i += stride) dst[i] = (D)src[i];
}

template <typename S>
static int convert_from(bcmi_queue q, int dst_dtype, const void *src, void *dst, size_t n)
{
    const unsigned blocks = (unsigned)std::min<size_t>((n + kTfThreads - 1) / kTfThreads, (size_t)cu_blocks(q, 8));
#define BCMI_CV(CODE, D)                                                                                          \
    case CODE:                                                                                                    \
        BCMI_LAUNCH(q, "convert", hipLaunchKernelGGL((k_convert<S, D>), dim3(blocks), dim3(kTfThreads), 0,        \
                                                     q->stream, (const S *)src, (D *)dst, n));                    \
        return 0
    switch (dst_dtype) {
        BCMI_CV(BCMI_I8, int8_t);
        BCMI_CV(BCMI_U8, uint8_t);
        BCMI_CV(BCMI_I16, int16_t);
        BCMI_CV(BCMI_U16, uint16_t);
        BCMI_CV(BCMI_I32, int32_t);
        BCMI_CV(BCMI_U32, uint32_t);
        BCMI_CV(BCMI_I64, int64_t);
        BCMI_CV(BCMI_U64, uint64_t);
        BCMI_CV(BCMI_F32, float);
        BCMI_CV(BCMI_F64, double);
    default: set_error("convert: bad destination dtype %d", dst_dtype); return BCMI_INVALID_VALUE;
    }
#undef BCMI_CV
}

static int check_iter(const bcmi_iter *it, bool allow_length)
{
    BCMI_CHECK(it->kind >= BCMI_ITER_BUFFER && it->kind <= BCMI_ITER_CONSTANT, BCMI_INVALID_VALUE,
               "transform: bad iterator kind %d", it->kind);
    BCMI_CHECK((it->fn >= BCMI_MAP_ID && it->fn <= (allow_length ? BCMI_MAP_LENGTH : BCMI_MAP_SQUARE)) ||
                   (it->fn >= BCMI_MAP_CMP && it->fn <= BCMI_MAP_CMP + BCMI_CMP_NE),
               BCMI_INVALID_VALUE, "transform: bad element map %d", it->fn);
    BCMI_CHECK(it->kind != BCMI_ITER_BUFFER || it->ptr, BCMI_INVALID_VALUE, "transform: null input array");
    return 0;
}

extern "C" int bcmi_transform_iter(bcmi_queue q, int dtype, int width, int op, const bcmi_iter *a,
                                   const bcmi_iter *b, void *out, size_t n)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(a, BCMI_INVALID_VALUE, "transform: null input descriptor");
    BCMI_TRY(check_iter(a, width > 1));
    if (b) BCMI_TRY(check_iter(b, false));
    if (n == 0) return 0;
    BCMI_CHECK(out, BCMI_INVALID_VALUE, "transform: null output");
    if (width > 1) {
        BCMI_CHECK((width == 2 || width == 4) && dtype == BCMI_F32 && !b && a->kind == BCMI_ITER_BUFFER &&
                       a->fn == BCMI_MAP_LENGTH,
                   BCMI_INVALID_VALUE, "transform: vector elements take length<floatN_> over an array only");
        const unsigned blocks =
            (unsigned)std::min<size_t>((n + kTfThreads - 1) / kTfThreads, (size_t)cu_blocks(q, 8));
        if (width == 2)
            BCMI_LAUNCH(q, "vector_length", hipLaunchKernelGGL((k_vector_length<2>), dim3(blocks), dim3(kTfThreads),
                                                                0, q->stream, (const float *)a->ptr, (float *)out, n));
        else
            BCMI_LAUNCH(q, "vector_length", hipLaunchKernelGGL((k_vector_length<4>), dim3(blocks), dim3(kTfThreads),
                                                                0, q->stream, (const float *)a->ptr, (float *)out, n));
        return 0;
    }
    BCMI_CHECK(width == 1, BCMI_INVALID_VALUE, "transform: element width %d unsupported", width);
    BCMI_CHECK(a->fn != BCMI_MAP_LENGTH, BCMI_INVALID_VALUE, "transform: length needs vector elements");
    switch (dtype) {
    case BCMI_I32: return transform_iter_typed<int32_t>(q, op, a, b, out, n);
    case BCMI_U32: return transform_iter_typed<uint32_t>(q, op, a, b, out, n);
    case BCMI_I64: return transform_iter_typed<int64_t>(q, op, a, b, out, n);
    case BCMI_U64: return transform_iter_typed<uint64_t>(q, op, a, b, out, n);
    case BCMI_F32: return transform_iter_typed<float>(q, op, a, b, out, n);
    case BCMI_F64: return transform_iter_typed<double>(q, op, a, b, out, n);
    default: set_error("transform: iterator inputs of dtype %d unsupported (32/64-bit int, float, double)", dtype);
             return BCMI_INVALID_VALUE;
    }
}

extern "C" int bcmi_convert(bcmi_queue q, int src_dtype, int dst_dtype, const void *src, void *dst, size_t n)
{
    BCMI_TRY(use_queue(q));
    if (n == 0) return 0;
    BCMI_CHECK(src && dst, BCMI_INVALID_VALUE, "convert: null pointer");
    switch (src_dtype) {
    case BCMI_I8: return convert_from<int8_t>(q, dst_dtype, src, dst, n);
    case BCMI_U8: return convert_from<uint8_t>(q, dst_dtype, src, dst, n);
    case BCMI_I16: return convert_from<int16_t>(q, dst_dtype, src, dst, n);
    case BCMI_U16: return convert_from<uint16_t>(q, dst_dtype, src, dst, n);
    case BCMI_I32: return convert_from<int32_t>(q, dst_dtype, src, dst, n);
    case BCMI_U32: return convert_from<uint32_t>(q, dst_dtype, src, dst, n);
    case BCMI_I64: return convert_from<int64_t>(q, dst_dtype, src, dst, n);
    case BCMI_U64: return convert_from<uint64_t>(q, dst_dtype, src, dst, n);
    case BCMI_F32: return convert_from<float>(q, dst_dtype, src, dst, n);
    case BCMI_F64: return convert_from<double>(q, dst_dtype, src, dst, n);
    default: set_error("convert: bad source dtype %d", src_dtype); return BCMI_INVALID_VALUE;
    }
}
