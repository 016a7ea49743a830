//---------------------------------------------------------------------------//
// sort (reference: algorithm/sort.hpp:182-202).  GPU column: n <= 32 the
// one-work-item insertion sort, otherwise radix (less<T> ascending,
// greater<T> descending); a user comparator (a comparator expression of
// lambda.hpp or experimental::less_by) takes the comparison sort
// (detail/merge_sort_on_gpu.hpp).  Host iterators are copied to the device, sorted
// and copied back (the reference maps them, sort.hpp:124-148).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_SORT_HPP
#define BOOST_COMPUTE_ALGORITHM_SORT_HPP

#include <iterator>

#include <boost/compute/algorithm/detail/merge_sort_on_gpu.hpp>
#include <boost/compute/algorithm/detail/radix_sort.hpp>
#include <boost/compute/container/vector.hpp>
#include <boost/compute/functional/operator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T, class Compare>
inline void dispatch_sort(int api, buffer_iterator<T> first, buffer_iterator<T> last, Compare compare,
                          command_queue &queue, std::true_type)
{
    sort_call(api, first, last, (T *)0, compare, queue);
}

template <class Iterator, class Compare>
inline void dispatch_sort(int api, Iterator first, Iterator last, Compare compare, command_queue &queue,
                          std::false_type)
{
    typedef typename std::iterator_traits<Iterator>::value_type T;
    ::boost::compute::vector<T> tmp(first, last, queue);
    dispatch_sort(api, tmp.begin(), tmp.end(), compare, queue, std::true_type());
    ::boost::compute::copy(tmp.begin(), tmp.end(), first, queue);
}

}  // namespace detail

template <class Iterator, class Compare>
inline void sort(Iterator first, Iterator last, Compare compare, command_queue &queue = system::default_queue())
{
    detail::dispatch_sort(BCMI_SORT, first, last, compare, queue,
                          std::integral_constant<bool, is_device_iterator<Iterator>::value>());
}

template <class Iterator>
inline void sort(Iterator first, Iterator last, command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<Iterator>::value_type T;
    ::boost::compute::sort(first, last, ::boost::compute::less<T>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
