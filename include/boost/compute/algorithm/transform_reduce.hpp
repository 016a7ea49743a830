//---------------------------------------------------------------------------//
// transform_reduce (reference: algorithm/transform_reduce.hpp): reduce of a
// mapped range, one pass, no temporary.  result is a device iterator or a
// host pointer (then blocking), as for reduce().  An empty range leaves
// result untouched.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_TRANSFORM_REDUCE_HPP
#define BOOST_COMPUTE_ALGORITHM_TRANSFORM_REDUCE_HPP

#include <iterator>

#include <boost/compute/algorithm/reduce.hpp>
#include <boost/compute/algorithm/transform.hpp>
#include <boost/compute/detail/mapreduce.hpp>
#include <boost/compute/functional/geometry.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T>
inline void store_result(const bcmi_seq &a, const bcmi_seq *b, size_t n, const bcmi_map &m, int rop,
                         buffer_iterator<T> result, command_queue &queue)
{
    run_map_reduce<T>(a, b, n, m, rop, (const T *)0, result.device_ptr(), queue);
}

template <class T>
inline void store_result(const bcmi_seq &a, const bcmi_seq *b, size_t n, const bcmi_map &m, int rop, T *result,
                         command_queue &queue)
{
    *result = map_reduce_value<T>(a, b, n, m, rop, (const T *)0, queue);
}

}  // namespace detail

template <class InputIterator, class OutputIterator, class UnaryTransformFunction, class BinaryReduceFunction>
inline void transform_reduce(InputIterator first, InputIterator last, OutputIterator result,
                             UnaryTransformFunction transform_function, BinaryReduceFunction,
                             command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "transform_reduce needs device iterators");
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    const size_t n = (size_t)(last - first);
    if (n == 0) return;
    const bcmi_map m = detail::unary_map_of<T, UnaryTransformFunction>::get(transform_function);
    detail::store_result<T>(detail::seq_of(first), (const bcmi_seq *)0, n, m, BinaryReduceFunction::bcmi_op, result,
                            queue);
}

// length<floatN_>: the lengths written out (one launch), then reduced
// (test/test_transform_reduce.cpp:42-60)
template <class S, std::size_t N, class OutputIterator, class BinaryReduceFunction>
inline void transform_reduce(buffer_iterator<vector_type<S, N> > first, buffer_iterator<vector_type<S, N> > last,
                             OutputIterator result, length<vector_type<S, N> > f, BinaryReduceFunction op,
                             command_queue &queue = system::default_queue())
{
    const size_t n = (size_t)(last - first);
    if (n == 0) return;
    ::boost::compute::vector<S> lengths(n, queue.get_context());
    ::boost::compute::transform(first, last, lengths.begin(), f, queue);
    ::boost::compute::reduce(lengths.begin(), lengths.end(), result, op, queue);
    queue.finish();  // lengths is released on return
}

template <class InputIterator1, class InputIterator2, class OutputIterator, class BinaryTransformFunction,
          class BinaryReduceFunction>
inline void transform_reduce(InputIterator1 first1, InputIterator1 last1, InputIterator2 first2,
                             OutputIterator result, BinaryTransformFunction transform_function, BinaryReduceFunction,
                             command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator1>::value, "transform_reduce needs device iterators");
    static_assert(is_device_iterator<InputIterator2>::value, "transform_reduce needs device iterators");
    typedef typename std::iterator_traits<InputIterator1>::value_type T;
    const size_t n = (size_t)(last1 - first1);
    if (n == 0) return;
    const bcmi_map m = detail::binary_map_of<T, BinaryTransformFunction>::get(transform_function);
    const bcmi_seq b = detail::seq_of(first2);
    detail::store_result<T>(detail::seq_of(first1), &b, n, m, BinaryReduceFunction::bcmi_op, result, queue);
}

}  // namespace compute
}  // namespace boost

#endif
