//---------------------------------------------------------------------------//
// Maps the inputs and function objects of transform_reduce / inner_product
// to the C-ABI's bcmi_seq / bcmi_map: a buffer_iterator is an array, a
// counting_iterator a generated sequence; abs<T> and pow(_1 [- c], 2) are
// unary maps, plus / minus / multiplies<T> binary ones.  Anything else does
// not compile.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_DETAIL_MAPREDUCE_HPP
#define BOOST_COMPUTE_DETAIL_MAPREDUCE_HPP

#include <cstdint>
#include <cstring>
#include <type_traits>

#include <bcmi.h>

#include <boost/compute/container/vector.hpp>
#include <boost/compute/functional/integer.hpp>
#include <boost/compute/functional/operator.hpp>
#include <boost/compute/iterator/buffer_iterator.hpp>
#include <boost/compute/iterator/counting_iterator.hpp>
#include <boost/compute/lambda.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T>
inline bcmi_seq seq_of(const buffer_iterator<T> &it)
{
    bcmi_seq s;
    s.ptr = it.device_ptr();
    s.start = 0;
    return s;
}

template <class T>
inline bcmi_seq seq_of(const counting_iterator<T> &it)
{
    static_assert(std::is_integral<T>::value, "counting_iterator inputs: integer element types");
    bcmi_seq s;
    s.ptr = NULL;
    s.start = (int64_t)*it;
    return s;
}

inline bcmi_map make_map(int op, int fn, uint64_t offset)
{
    bcmi_map m;
    m.op = op;
    m.fn = fn;
    m.offset = offset;
    return m;
}

template <class T, class F>
struct unary_map_of {
    static_assert(sizeof(F) == 0, "transform_reduce: unary function not recognised (abs<T>, pow(_1, 2), "
                                  "pow(_1 - c, 2))");
};

template <class T, class U>
struct unary_map_of<T, ::boost::compute::abs<U> > {
    static bcmi_map get(const ::boost::compute::abs<U> &) { return make_map(0, BCMI_MAP_ABS, 0); }
};

template <class T, class C>
struct unary_map_of<T, lambda::square_expr<C> > {
    // OpenCL's pow is for floats; an integer offset converts exactly as C
    // would, a float offset must have the element type
    static_assert(std::is_floating_point<T>::value, "pow(_1 - c, 2) needs float elements");
    static_assert(std::is_integral<C>::value || std::is_same<C, T>::value,
                  "pow(_1 - c, 2): c must be an integer or have the element type");
    static bcmi_map get(const lambda::square_expr<C> &e)
    {
        const T c = static_cast<T>(e.c);
        uint64_t bits = 0;
        std::memcpy(&bits, &c, sizeof(T));
        return make_map(0, BCMI_MAP_SQUARE, bits);
    }
};

template <class T, class F>
struct binary_map_of {
    static_assert(sizeof(F) == 0, "binary transform not recognised (plus, minus, multiplies)");
};
template <class T, class U>
struct binary_map_of<T, plus<U> > {
    static bcmi_map get(const plus<U> &) { return make_map(BCMI_PLUS, BCMI_MAP_ID, 0); }
};
template <class T, class U>
struct binary_map_of<T, minus<U> > {
    static bcmi_map get(const minus<U> &) { return make_map(BCMI_MINUS, BCMI_MAP_ID, 0); }
};
template <class T, class U>
struct binary_map_of<T, multiplies<U> > {
    static bcmi_map get(const multiplies<U> &) { return make_map(BCMI_MULTIPLIES, BCMI_MAP_ID, 0); }
};

// one value reduced into *out_device (blocking read when the result goes to
// the host)
template <class T>
inline void run_map_reduce(const bcmi_seq &a, const bcmi_seq *b, size_t n, const bcmi_map &m, int rop, const T *init,
                           T *out_device, command_queue &queue)
{
    bcmi_check(bcmi_transform_reduce(queue.get(), bcmi_dtype<T>(), &a, b, n, &m, rop, init, out_device, 0,
                                     BCMI_PRECISION_COMPENSATED));
}

template <class T>
inline T map_reduce_value(const bcmi_seq &a, const bcmi_seq *b, size_t n, const bcmi_map &m, int rop, const T *init,
                          command_queue &queue)
{
    ::boost::compute::vector<T> tmp(1, queue.get_context());
    run_map_reduce(a, b, n, m, rop, init, tmp.begin().device_ptr(), queue);
    return tmp.begin().read(queue);
}

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif
