// ops.hpp -- the binary functors the dispatch recognises, as device code.
//
// plus<T>, multiplies<T>, min<T>, max<T> (functional/operator.hpp:57-96,
// functional/common.hpp).  Integer arithmetic wraps modulo 2^W as in OpenCL
// C, so it is done on the unsigned type.
#pragma once

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
BHD T apply_op(T a, T b)
{
    using U = typename wrap_t<T>::type;
    if constexpr (OP == BCMI_PLUS) return (T)(U)((U)a + (U)b);
    else if constexpr (OP == BCMI_MULTIPLIES) return (T)(U)((U)a * (U)b);
    else if constexpr (OP == BCMI_MIN) return b < a ? b : a;
    else return a < b ? b : a;
}

template <typename T, int OP>
BHD T op_identity()
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

}  // namespace bcmi
