//---------------------------------------------------------------------------//
// stable_sort (reference: algorithm/stable_sort.hpp:81-110): radix at any n;
// a user comparator takes the comparison sort (stable_sort.hpp:36-50).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_STABLE_SORT_HPP
#define BOOST_COMPUTE_ALGORITHM_STABLE_SORT_HPP

#include <boost/compute/algorithm/sort.hpp>

namespace boost {
namespace compute {

template <class Iterator, class Compare>
inline void stable_sort(Iterator first, Iterator last, Compare compare,
                        command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<Iterator>::value, "stable_sort needs device iterators");
    detail::dispatch_sort(BCMI_STABLE_SORT, first, last, compare, queue, std::true_type());
}

template <class Iterator>
inline void stable_sort(Iterator first, Iterator last, command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<Iterator>::value_type T;
    ::boost::compute::stable_sort(first, last, ::boost::compute::less<T>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
