//---------------------------------------------------------------------------//
// inner_product (reference: algorithm/inner_product.hpp): init plus the sum
// of products, or, in the general form, init accumulate_function'd with
// transform_function(a[i], b[i]), in one pass without the reference's
// temporary vector.  Float sums are compensated and parallel (the reference
// accumulates floats in one work-item).  Blocking: returns the value.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_INNER_PRODUCT_HPP
#define BOOST_COMPUTE_ALGORITHM_INNER_PRODUCT_HPP

#include <iterator>
#include <type_traits>

#include <boost/compute/detail/mapreduce.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class InputIterator1, class InputIterator2, class T, class BinaryAccumulateFunction,
          class BinaryTransformFunction>
inline T inner_product(InputIterator1 first1, InputIterator1 last1, InputIterator2 first2, T init,
                       BinaryAccumulateFunction, BinaryTransformFunction transform_function,
                       command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator1>::value, "inner_product needs device iterators");
    static_assert(is_device_iterator<InputIterator2>::value, "inner_product needs device iterators");
    typedef typename std::iterator_traits<InputIterator1>::value_type V;
    static_assert(std::is_same<T, V>::value, "inner_product: init must have the element type");
    const size_t n = (size_t)(last1 - first1);
    if (n == 0) return init;
    const bcmi_map m = detail::binary_map_of<V, BinaryTransformFunction>::get(transform_function);
    const bcmi_seq b = detail::seq_of(first2);
    return detail::map_reduce_value<V>(detail::seq_of(first1), &b, n, m, BinaryAccumulateFunction::bcmi_op, &init,
                                       queue);
}

template <class InputIterator1, class InputIterator2, class T>
inline T inner_product(InputIterator1 first1, InputIterator1 last1, InputIterator2 first2, T init,
                       command_queue &queue = system::default_queue())
{
    return ::boost::compute::inner_product(first1, last1, first2, init, plus<T>(), multiplies<T>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
