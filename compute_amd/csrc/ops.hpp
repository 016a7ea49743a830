// ops.hpp -- the binary functors the dispatch recognises, as device code.
//
// plus<T>, multiplies<T>, min<T>, max<T> (functional/operator.hpp:57-96,
// functional/common.hpp).  Integer arithmetic wraps modulo 2^W as in OpenCL
// C, so it is done on the unsigned type.
#pragma once

#include <cstring>
#include <limits>
#include <type_traits>

#include "device.hpp"

namespace bcmi {

template <typename T, bool = std::is_integral<T>::value>
struct wrap_t { using type = T; };
template <typename T>
struct wrap_t<T, true> {
    // at least 32 bits so that narrow products cannot overflow int
    using type = std::conditional_t<(sizeof(T) < 4), uint32_t, std::make_unsigned_t<T>>;
};

template <typename T, int OP>
struct OpImpl {
    static BHD T apply(T a, T b)
    {
        using U = typename wrap_t<T>::type;
        if constexpr (OP == BCMI_PLUS) return (T)(U)((U)a + (U)b);
        else if constexpr (OP == BCMI_MULTIPLIES) return (T)(U)((U)a * (U)b);
        else if constexpr (OP == BCMI_MIN) return b < a ? b : a;
        else return a < b ? b : a;
    }
    static BHD T id()
    {
        if constexpr (OP == BCMI_PLUS) return (T)0;
        else if constexpr (OP == BCMI_MULTIPLIES) return (T)1;
        else if constexpr (OP == BCMI_MIN) {
            if constexpr (std::is_floating_point<T>::value) return __builtin_inf();
            else return std::numeric_limits<T>::max();
        } else {
            if constexpr (std::is_floating_point<T>::value) return -__builtin_inf();
            else return std::numeric_limits<T>::lowest();
        }
    }
};

// An OpenCL vector element (int2_, float4_ ...): W components of E, the
// functors applied component by component (functional/operator.hpp on
// vector types).
template <typename E, int W>
struct vecn {
    E c[W];
};

template <typename E, int W, int OP>
struct OpImpl<vecn<E, W>, OP> {
    static BHD vecn<E, W> apply(vecn<E, W> a, vecn<E, W> b)
    {
        vecn<E, W> r;
        for (int i = 0; i < W; ++i) r.c[i] = OpImpl<E, OP>::apply(a.c[i], b.c[i]);
        return r;
    }
    static BHD vecn<E, W> id()
    {
        vecn<E, W> r;
        for (int i = 0; i < W; ++i) r.c[i] = OpImpl<E, OP>::id();
        return r;
    }
};

template <typename T, int OP>
BHD T apply_op(T a, T b)
{
    return OpImpl<T, OP>::apply(a, b);
}

template <typename T, int OP>
BHD T op_identity()
{
    return OpImpl<T, OP>::id();
}

template <typename T>
BHD T map_bits(uint64_t b)
{
    T v;
    memcpy(&v, &b, sizeof(T));
    return v;
}

template <typename T>
BHD T map_abs(T x)
{
    if constexpr (std::is_floating_point<T>::value) {
        return __builtin_fabs(x);
    } else if constexpr (std::is_signed<T>::value) {
        using U = typename wrap_t<T>::type;
        return x < 0 ? (T)(U)(0u - (U)x) : x;  // abs(INT_MIN) wraps to INT_MIN, as OpenCL's abs bits
    } else {
        return x;
    }
}

template <typename T>
BHD T map_sub(T a, T b)
{
    using U = typename wrap_t<T>::type;
    if constexpr (std::is_floating_point<T>::value) return a - b;
    else return (T)(U)((U)a - (U)b);
}

// An input sequence as the reference's iterators give it: an array
// (buffer_iterator), value + i (counting_iterator) or value
// (constant_iterator), then y = g(element - offset) for the element map g
// of a transform_iterator (identity, abs, square), or a comparison with
// offset giving 1 or 0.  kind and fn are run-time values that are uniform
// across a launch.
template <typename T>
struct IterIn {
    const T *ptr;
    T value, offset;
    int32_t kind, fn;
    BHD bool array() const { return kind == BCMI_ITER_BUFFER; }
    BHD T gen(size_t i) const
    {
        return kind == BCMI_ITER_COUNTING ? apply_op<T, BCMI_PLUS>(value, (T)i) : value;
    }
    BHD T map(T x) const
    {
        if (fn >= BCMI_MAP_CMP) {
            bool r;
            switch (fn - BCMI_MAP_CMP) {
            case BCMI_CMP_LT: r = x < offset; break;
            case BCMI_CMP_LE: r = x <= offset; break;
            case BCMI_CMP_GT: r = x > offset; break;
            case BCMI_CMP_GE: r = x >= offset; break;
            case BCMI_CMP_EQ: r = x == offset; break;
            default: r = x != offset; break;
            }
            return r ? T(1) : T(0);
        }
        const T t = map_sub(x, offset);
        if (fn == BCMI_MAP_ABS) return map_abs(t);
        if (fn == BCMI_MAP_SQUARE) return apply_op<T, BCMI_MULTIPLIES>(t, t);
        return t;
    }
    BHD T at(size_t i) const { return map(array() ? ptr[i] : gen(i)); }
};

// The kernels' view of a C-ABI bcmi_iter whose elements are of type T.
template <typename T>
inline IterIn<T> iter_in(const bcmi_iter *in)
{
    IterIn<T> s;
    s.ptr = (const T *)in->ptr;
    s.value = map_bits<T>(in->value);
    s.offset = map_bits<T>(in->offset);
    s.kind = in->kind;
    s.fn = in->fn;
    return s;
}

}  // namespace bcmi
