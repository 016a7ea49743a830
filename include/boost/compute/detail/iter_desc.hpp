//---------------------------------------------------------------------------//
// The C-ABI view (bcmi_iter) of an input iterator: a buffer_iterator is an
// array, a counting_iterator / constant_iterator a generated sequence, and
// a transform_iterator over one of those adds the element map of its
// function.  A function the kernels do not implement does not compile.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_DETAIL_ITER_DESC_HPP
#define BOOST_COMPUTE_DETAIL_ITER_DESC_HPP

#include <cstdint>
#include <cstring>
#include <type_traits>

#include <bcmi.h>

#include <boost/compute/detail/bcmi.hpp>
#include <boost/compute/functional/integer.hpp>
#include <boost/compute/iterator/buffer_iterator.hpp>
#include <boost/compute/iterator/constant_iterator.hpp>
#include <boost/compute/iterator/counting_iterator.hpp>
#include <boost/compute/iterator/transform_iterator.hpp>
#include <boost/compute/lambda.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T>
inline uint64_t element_bits(T v)
{
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}

inline bcmi_iter make_iter(const void *ptr, int kind, uint64_t value)
{
    bcmi_iter d;
    d.ptr = ptr;
    d.kind = kind;
    d.fn = BCMI_MAP_ID;
    d.value = value;
    d.offset = 0;
    return d;
}

template <class It>
struct iter_desc {
    static_assert(sizeof(It) == 0, "input iterator not supported (buffer_iterator, counting_iterator, "
                                   "constant_iterator, transform_iterator over one of those)");
};

template <class T>
struct iter_desc<buffer_iterator<T> > {
    typedef T value_type;
    static bcmi_iter get(const buffer_iterator<T> &it) { return make_iter(it.device_ptr(), BCMI_ITER_BUFFER, 0); }
};

template <class T>
struct iter_desc<counting_iterator<T> > {
    typedef T value_type;
    static bcmi_iter get(const counting_iterator<T> &it) { return make_iter(0, BCMI_ITER_COUNTING, element_bits(*it)); }
};

template <class T>
struct iter_desc<constant_iterator<T> > {
    typedef T value_type;
    static bcmi_iter get(const constant_iterator<T> &it)
    {
        return make_iter(0, BCMI_ITER_CONSTANT, element_bits(it.value()));
    }
};

// element map of a transform_iterator's function on elements of type T
template <class T, class F>
struct element_map_of {
    static_assert(sizeof(F) == 0, "transform_iterator function not recognised (_1 - c, _1 + c, abs<T>(), "
                                  "_1 * _1, pow(_1, 2), pown(_1, 2), pow(_1 - c, 2), _1 cmp c)");
};

template <class T, class C>
struct element_map_of<T, lambda::minus_arg1<C> > {
    static void apply(const lambda::minus_arg1<C> &f, bcmi_iter &d) { d.offset = element_bits((T)f.c); }
};

template <class T, class C>
struct element_map_of<T, lambda::plus_arg1<C> > {
    // element + c == element - (-c), exactly for floats and modulo 2^W for
    // integers
    static void apply(const lambda::plus_arg1<C> &f, bcmi_iter &d)
    {
        d.offset = element_bits((T)(T(0) - (T)f.c));
    }
};

template <class T, class U>
struct element_map_of<T, ::boost::compute::abs<U> > {
    static void apply(const ::boost::compute::abs<U> &, bcmi_iter &d) { d.fn = BCMI_MAP_ABS; }
};

template <class T, class C>
struct element_map_of<T, lambda::square_expr<C> > {
    static void apply(const lambda::square_expr<C> &f, bcmi_iter &d)
    {
        d.fn = BCMI_MAP_SQUARE;
        d.offset = element_bits((T)f.c);
    }
};

template <class T, class L>
struct element_map_of<T, lambda::pred_expr<BCMI_FN_ID, int, L> > {
    static void apply(const lambda::pred_expr<BCMI_FN_ID, int, L> &f, bcmi_iter &d)
    {
        d.fn = BCMI_MAP_CMP + f.cmp;
        d.offset = element_bits((T)f.c);
    }
};

template <class It, class F>
struct iter_desc<transform_iterator<It, F> > {
    typedef typename iter_desc<It>::value_type value_type;
    static bcmi_iter get(const transform_iterator<It, F> &it)
    {
        static_assert(std::is_same<It, buffer_iterator<value_type> >::value ||
                          std::is_same<It, counting_iterator<value_type> >::value ||
                          std::is_same<It, constant_iterator<value_type> >::value,
                      "transform_iterator over a buffer, counting or constant iterator (one map)");
        bcmi_iter d = iter_desc<It>::get(it.base());
        element_map_of<value_type, F>::apply(it.functor(), d);
        return d;
    }
};

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif
