//---------------------------------------------------------------------------//
// Maps a recognised unary predicate over elements of type T to the C-ABI's
// bcmi_predicate.  The reference evaluates the printed OpenCL C expression
// with C's usual arithmetic conversions; the kernels compare in T promoted
// (int for 8/16-bit integers), so a literal whose conversion would give a
// different comparison type is rejected at compile time, except a double
// literal against float elements, accepted when it is exactly a float (then
// both comparisons agree).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_DETAIL_PREDICATE_HPP
#define BOOST_COMPUTE_DETAIL_PREDICATE_HPP

#include <cstdint>
#include <cstring>
#include <type_traits>

#include <bcmi.h>

#include <boost/compute/detail/bcmi.hpp>
#include <boost/compute/functional/logical.hpp>
#include <boost/compute/functional/relational.hpp>
#include <boost/compute/lambda.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T>
struct promoted {
    typedef typename std::conditional<std::is_integral<T>::value && (sizeof(T) < 4), int, T>::type type;
};

template <class C>
inline uint64_t value_bits(C v)
{
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(C));
    return b;
}

// right-hand side of the comparison
template <class T, class L>
inline uint64_t comparison_bits(const L &v)
{
    typedef typename promoted<T>::type C;
    typedef typename std::common_type<C, L>::type CT;
    const bool float_vs_double = std::is_same<C, float>::value && std::is_same<L, double>::value;
    static_assert(std::is_same<CT, C>::value || (std::is_same<C, float>::value && std::is_same<L, double>::value),
                  "predicate literal: its type would change the comparison type; use a literal of the element type");
    const C c = static_cast<C>(v);
    if (float_vs_double && static_cast<double>(c) != static_cast<double>(v))
        throw opencl_error(BCMI_INVALID_VALUE, "predicate literal: a double that is not exactly a float");
    return value_bits(c);
}

// operand of _1 * a, _1 % a, _1 & a
template <class T, class A>
inline uint64_t operand_bits(const A &a)
{
    typedef typename promoted<T>::type C;
    static_assert(std::is_same<typename std::common_type<C, A>::type, C>::value,
                  "predicate operand: its type would change the arithmetic type; use a literal of the element type");
    return value_bits(static_cast<C>(a));
}

template <class T, class P>
struct predicate_of {
    static_assert(sizeof(P) == 0,
                  "predicate not recognised: use f(_1) cmp c with f = _1, _1*a, _1%a, _1&a, signbit_<T>() or not1()");
};

template <class T, int FN, class A, class L>
struct predicate_of<T, lambda::pred_expr<FN, A, L> > {
    static bcmi_predicate get(const lambda::pred_expr<FN, A, L> &e)
    {
        bcmi_predicate p;
        p.fn = FN;
        p.cmp = e.cmp;
        p.negate = 0;
        p.reserved = 0;
        p.operand = FN == BCMI_FN_ID ? 0 : operand_bits<T>(e.a);
        p.value = comparison_bits<T>(e.c);
        return p;
    }
};

// f(_1) alone (e.g. _1 & 1): true where it is nonzero, as C reads a condition
template <class T, int FN, class A>
struct predicate_of<T, lambda::arg1_expr<FN, A> > {
    static bcmi_predicate get(const lambda::arg1_expr<FN, A> &e)
    {
        bcmi_predicate p;
        p.fn = FN;
        p.cmp = BCMI_CMP_NE;
        p.negate = 0;
        p.reserved = 0;
        p.operand = operand_bits<T>(e.a);
        p.value = 0;
        return p;
    }
};

template <class T, class U>
struct predicate_of<T, signbit_<U> > {
    static_assert(std::is_floating_point<T>::value, "signbit_ needs float elements");
    static bcmi_predicate get(const signbit_<U> &)
    {
        bcmi_predicate p;
        p.fn = BCMI_FN_SIGNBIT;
        p.cmp = BCMI_CMP_NE;
        p.negate = 0;
        p.reserved = 0;
        p.operand = 0;
        p.value = 0;
        return p;
    }
};

template <class T, class P>
struct predicate_of<T, unary_negate<P> > {
    static bcmi_predicate get(const unary_negate<P> &n)
    {
        bcmi_predicate p = predicate_of<T, P>::get(n.predicate());
        p.negate = !p.negate;
        return p;
    }
};

template <class T, class P>
inline bcmi_predicate make_predicate(const P &pred)
{
    return predicate_of<T, P>::get(pred);
}

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif
