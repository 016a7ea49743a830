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
//                                     (lambda/functional.hpp:521) and of
//   _1 + c, _1 - c, _1 * _1,          transform_iterator (scan, reduce,
//   pown(_1, 2), _1 cmp c             transform, copy)
//   _1 < _2, get<N>(_1) < get<M>(_2), abs(_1) < abs(_2), length(_1) <
//   length(_2), abs(get<N>(_1) + get<M>(_1)) < ... (and >)
//                                     comparators of the comparison sort
// Anything else does not compile.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_LAMBDA_HPP
#define BOOST_COMPUTE_LAMBDA_HPP

#include <cstddef>
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

// _1 + c, _1 * _1 and pown(_1, 2): more element maps of transform_iterator
// (test/test_reduce.cpp:216-234, test/test_scan.cpp:273-276)
template <class T>
struct plus_arg1 {
    T c;
};

template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
inline plus_arg1<T> operator+(placeholder<1>, T c)
{
    return plus_arg1<T>{c};
}

inline square_expr<int> operator*(placeholder<1>, placeholder<1>)
{
    return square_expr<int>{0};
}

inline square_expr<int> pown(placeholder<1>, int e)
{
    if (e != 2) throw opencl_error(BCMI_INVALID_VALUE, "pown: only the exponent 2 is supported");
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

// Comparator forms: the user comparators of sort / stable_sort /
// sort_by_key that the reference's tests write as OpenCL functions
// (test/test_merge_sort_gpu.cpp:239-242, 275-278, 311-314; test/test_sort.cpp:260;
// test/test_stable_sort.cpp:53-56):
//   _1 < _2, _1 > _2                       the element (a scalar)
//   get<N>(_1) < get<M>(_2)                components (M may differ from N)
//   abs(_1) < abs(_2), abs(get<N>(_1)) ... unsigned magnitude of integers
//   abs(get<N>(_1) + get<M>(_1)) < abs(get<N>(_2) + get<M>(_2))
//   length(_1) < length(_2)                Euclidean length of a float vector
// and the same with >.
template <int I, int FORM, std::size_t C0, std::size_t C1 = 0>
struct key_term {
};

template <int I, std::size_t C0, std::size_t C1>
struct sum_term {
};

template <int FORM, std::size_t LHS, std::size_t RHS, std::size_t SECOND, bool GREATER>
struct compare_expr {
};

template <std::size_t N, int I>
inline key_term<I, BCMI_KEY_FIELD, N> get(placeholder<I>)
{
    return key_term<I, BCMI_KEY_FIELD, N>();
}

template <int I>
inline key_term<I, BCMI_KEY_ABS, 0> abs(placeholder<I>)
{
    return key_term<I, BCMI_KEY_ABS, 0>();
}

template <int I, std::size_t N>
inline key_term<I, BCMI_KEY_ABS, N> abs(key_term<I, BCMI_KEY_FIELD, N>)
{
    return key_term<I, BCMI_KEY_ABS, N>();
}

template <int I, std::size_t A, std::size_t B>
inline sum_term<I, A, B> operator+(key_term<I, BCMI_KEY_FIELD, A>, key_term<I, BCMI_KEY_FIELD, B>)
{
    return sum_term<I, A, B>();
}

template <int I, std::size_t A, std::size_t B>
inline key_term<I, BCMI_KEY_ABS_SUM2, A, B> abs(sum_term<I, A, B>)
{
    return key_term<I, BCMI_KEY_ABS_SUM2, A, B>();
}

template <int I>
inline key_term<I, BCMI_KEY_LENGTH, 0> length(placeholder<I>)
{
    return key_term<I, BCMI_KEY_LENGTH, 0>();
}

inline compare_expr<BCMI_KEY_FIELD, 0, 0, 0, false> operator<(placeholder<1>, placeholder<2>)
{
    return compare_expr<BCMI_KEY_FIELD, 0, 0, 0, false>();
}

inline compare_expr<BCMI_KEY_FIELD, 0, 0, 0, true> operator>(placeholder<1>, placeholder<2>)
{
    return compare_expr<BCMI_KEY_FIELD, 0, 0, 0, true>();
}

#define BOOST_COMPUTE_LAMBDA_KEY_CMP(OP, GREATER)                                                          \
    template <int F, std::size_t A, std::size_t B, std::size_t A2, std::size_t B2>                         \
    inline compare_expr<F, A, A2, B, GREATER> operator OP(key_term<1, F, A, B>, key_term<2, F, A2, B2>)    \
    {                                                                                                      \
        static_assert(F == BCMI_KEY_FIELD || F == BCMI_KEY_ABS || (A == A2 && B == B2),                    \
                      "abs-of-sum and length comparators read the same components of both operands");     \
        return compare_expr<F, A, A2, B, GREATER>();                                                       \
    }
BOOST_COMPUTE_LAMBDA_KEY_CMP(<, false)
BOOST_COMPUTE_LAMBDA_KEY_CMP(>, true)
#undef BOOST_COMPUTE_LAMBDA_KEY_CMP

}  // namespace lambda

using lambda::_1;
using lambda::_2;

}  // namespace compute
}  // namespace boost

#endif
