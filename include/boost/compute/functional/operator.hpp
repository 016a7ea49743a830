//---------------------------------------------------------------------------//
// Functor tags (reference: functional/operator.hpp:57-96,
// functional/common.hpp).  In this build they are tags the algorithms
// dispatch on to precompiled gfx950 kernels; each also works on the host.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_FUNCTIONAL_OPERATOR_HPP
#define BOOST_COMPUTE_FUNCTIONAL_OPERATOR_HPP

#include <bcmi.h>

namespace boost {
namespace compute {

#define BOOST_COMPUTE_BINARY_TAG(NAME, EXPR, OPCODE)                   \
    template <class T>                                                 \
    struct NAME {                                                      \
        typedef T result_type;                                         \
        static const int bcmi_op = OPCODE;                             \
        T operator()(const T &x, const T &y) const { return EXPR; }    \
    };

BOOST_COMPUTE_BINARY_TAG(plus, x + y, BCMI_PLUS)
BOOST_COMPUTE_BINARY_TAG(minus, x - y, BCMI_MINUS)
BOOST_COMPUTE_BINARY_TAG(multiplies, x * y, BCMI_MULTIPLIES)
BOOST_COMPUTE_BINARY_TAG(divides, x / y, BCMI_DIVIDES)
BOOST_COMPUTE_BINARY_TAG(min, y < x ? y : x, BCMI_MIN)
BOOST_COMPUTE_BINARY_TAG(max, x < y ? y : x, BCMI_MAX)
#undef BOOST_COMPUTE_BINARY_TAG

template <class T>
struct less {
    typedef bool result_type;
    bool operator()(const T &x, const T &y) const { return x < y; }
};

template <class T>
struct greater {
    typedef bool result_type;
    bool operator()(const T &x, const T &y) const { return x > y; }
};

template <class T>
struct equal_to {
    typedef bool result_type;
    bool operator()(const T &x, const T &y) const { return x == y; }
};

}  // namespace compute
}  // namespace boost

#endif
