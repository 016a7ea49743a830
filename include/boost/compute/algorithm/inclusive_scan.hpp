//---------------------------------------------------------------------------//
// inclusive_scan (reference: algorithm/inclusive_scan.hpp:53-87).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_INCLUSIVE_SCAN_HPP
#define BOOST_COMPUTE_ALGORITHM_INCLUSIVE_SCAN_HPP

#include <boost/compute/algorithm/detail/scan.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class OutputIterator, class BinaryOperator>
inline OutputIterator inclusive_scan(InputIterator first, InputIterator last, OutputIterator result,
                                     BinaryOperator binary_op, command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "inclusive_scan needs device iterators");
    static_assert(is_device_iterator<OutputIterator>::value, "inclusive_scan needs device iterators");
    typedef typename std::iterator_traits<OutputIterator>::value_type O;
    return detail::scan(first, last, result, false, O(0), binary_op, queue);
}

template <class InputIterator, class OutputIterator>
inline OutputIterator inclusive_scan(InputIterator first, InputIterator last, OutputIterator result,
                                     command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<OutputIterator>::value_type O;
    return ::boost::compute::inclusive_scan(first, last, result, ::boost::compute::plus<O>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
