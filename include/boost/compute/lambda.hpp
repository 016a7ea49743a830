//---------------------------------------------------------------------------//
// The lambda placeholders _1, _2 (reference: lambda/placeholders.hpp).  The
// reference prints arbitrary lambda expressions as OpenCL C
// (lambda/context.hpp:145-156); this build has no run-time compiler, so it
// recognises the expression forms its kernels implement and maps each to a
// precompiled kernel:
//   a * _1 + _2                       saxpy (perf/perf_saxpy.cpp:50-55)
//   f(_1) cmp c                       unary predicates of the selection
//                                     algorithms (copy_if, count_if, remove_if,
//                                     partition...), with f one of _1, _1 * a,
//                                     a * _1, _1 % a, _1 & a and cmp one of
//                                     < <= > >= == !=
//   pow(_1, 2), pow(_1 - c, 2)        element maps of transform_reduce
//                                     (lambda/functional.hpp:521)
// Anything else does not compile.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_LAMBDA_HPP
#define BOOST_COMPUTE_LAMBDA_HPP

#include <type_traits>

#include <bcmi.h>

#include <boost/compute/exception/opencl_error.hpp>

namespace boost {
namespace compute {
namespace lambda {

template <int I>
struct placeholder {
};

static const placeholder<1> _1 = placeholder<1>();
static const placeholder<2> _2 = placeholder<2>();

template <class T>
struct scaled_arg1 {
    T a;
};

template <class T>
struct saxpy_expr {
    T a;
};

template <class T>
inline scaled_arg1<T> operator*(T a, placeholder<1>)
{
    return scaled_arg1<T>{a};
}

template <class T>
inline scaled_arg1<T> operator*(placeholder<1>, T a)
{
    return scaled_arg1<T>{a};
}

template <class T>
inline saxpy_expr<T> operator+(scaled_arg1<T> s, placeholder<2>)
{
    return saxpy_expr<T>{s.a};
}

// f(_1) for f = _1 % a, _1 & a
template <int FN, class A>
struct arg1_expr {
    A a;
};

template <class A, class = typename std::enable_if<std::is_integral<A>::value>::type>
inline arg1_expr<BCMI_FN_MOD, A> operator%(placeholder<1>, A a)
{
    return arg1_expr<BCMI_FN_MOD, A>{a};
}

template <class A, class = typename std::enable_if<std::is_integral<A>::value>::type>
inline arg1_expr<BCMI_FN_AND, A> operator&(placeholder<1>, A a)
{
    return arg1_expr<BCMI_FN_AND, A>{a};
}

// _1 - c, and pow(_1 - c, 2) / pow(_1, 2): the squared-deviation element map
template <class T>
struct minus_arg1 {
    T c;
};

template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
inline minus_arg1<T> operator-(placeholder<1>, T c)
{
    return minus_arg1<T>{c};
}

template <class T>
struct square_expr {
    T c;  // offset subtracted before squaring
};

template <class T>
inline square_expr<T> pow(minus_arg1<T> m, int e)
{
    if (e != 2) throw opencl_error(BCMI_INVALID_VALUE, "pow: only the exponent 2 is supported");
    return square_expr<T>{m.c};
}

inline square_expr<int> pow(placeholder<1>, int e)
{
    if (e != 2) throw opencl_error(BCMI_INVALID_VALUE, "pow: only the exponent 2 is supported");
    return square_expr<int>{0};
}

// f(_1) cmp c
template <int FN, class A, class L>
struct pred_expr {
    A a;      // f's operand (unused for _1)
    L c;      // right-hand side
    int cmp;  // bcmi_cmp
};

#define BOOST_COMPUTE_LAMBDA_CMP(OP, CODE)                                                              \
    template <class L, class = typename std::enable_if<std::is_arithmetic<L>::value>::type>             \
    inline pred_expr<BCMI_FN_ID, int, L> operator OP(placeholder<1>, L c)                               \
    {                                                                                                   \
        return pred_expr<BCMI_FN_ID, int, L>{0, c, CODE};                                               \
    }                                                                                                   \
    template <class A, class L, class = typename std::enable_if<std::is_arithmetic<L>::value>::type>    \
    inline pred_expr<BCMI_FN_MUL, A, L> operator OP(scaled_arg1<A> s, L c)                              \
    {                                                                                                   \
        return pred_expr<BCMI_FN_MUL, A, L>{s.a, c, CODE};                                              \
    }                                                                                                   \
    template <int FN, class A, class L,                                                                 \
              class = typename std::enable_if<std::is_arithmetic<L>::value>::type>                      \
    inline pred_expr<FN, A, L> operator OP(arg1_expr<FN, A> f, L c)                                     \
    {                                                                                                   \
        return pred_expr<FN, A, L>{f.a, c, CODE};                                                       \
    }

BOOST_COMPUTE_LAMBDA_CMP(<, BCMI_CMP_LT)
BOOST_COMPUTE_LAMBDA_CMP(<=, BCMI_CMP_LE)
BOOST_COMPUTE_LAMBDA_CMP(>, BCMI_CMP_GT)
BOOST_COMPUTE_LAMBDA_CMP(>=, BCMI_CMP_GE)
BOOST_COMPUTE_LAMBDA_CMP(==, BCMI_CMP_EQ)
BOOST_COMPUTE_LAMBDA_CMP(!=, BCMI_CMP_NE)
#undef BOOST_COMPUTE_LAMBDA_CMP

}  // namespace lambda

using lambda::_1;
using lambda::_2;

}  // namespace compute
}  // namespace boost

#endif
