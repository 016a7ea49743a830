//---------------------------------------------------------------------------//
// exclusive_scan (reference: algorithm/exclusive_scan.hpp:55-104).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_EXCLUSIVE_SCAN_HPP
#define BOOST_COMPUTE_ALGORITHM_EXCLUSIVE_SCAN_HPP

#include <boost/compute/algorithm/detail/scan.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class OutputIterator, class T, class BinaryOperator>
inline OutputIterator exclusive_scan(InputIterator first, InputIterator last, OutputIterator result, T init,
                                     BinaryOperator binary_op, command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "exclusive_scan needs device iterators");
    static_assert(is_device_iterator<OutputIterator>::value, "exclusive_scan needs device iterators");
    return detail::scan(first, last, result, true, init, binary_op, queue);
}

template <class InputIterator, class OutputIterator, class T>
inline OutputIterator exclusive_scan(InputIterator first, InputIterator last, OutputIterator result, T init,
                                     command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<OutputIterator>::value_type O;
    return ::boost::compute::exclusive_scan(first, last, result, init, ::boost::compute::plus<O>(), queue);
}

template <class InputIterator, class OutputIterator>
inline OutputIterator exclusive_scan(InputIterator first, InputIterator last, OutputIterator result,
                                     command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<OutputIterator>::value_type O;
    return ::boost::compute::exclusive_scan(first, last, result, O(0), ::boost::compute::plus<O>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
