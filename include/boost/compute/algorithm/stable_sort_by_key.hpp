//---------------------------------------------------------------------------//
// stable_sort_by_key (reference: algorithm/stable_sort_by_key.hpp:135-163):
// n < 32 the insertion sort by key, otherwise radix; a user comparator takes
// the comparison sort (stable_sort_by_key.hpp:84-104).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_STABLE_SORT_BY_KEY_HPP
#define BOOST_COMPUTE_ALGORITHM_STABLE_SORT_BY_KEY_HPP

#include <boost/compute/algorithm/sort_by_key.hpp>

namespace boost {
namespace compute {

template <class KeyIterator, class ValueIterator, class Compare>
inline void stable_sort_by_key(KeyIterator keys_first, KeyIterator keys_last, ValueIterator values_first,
                               Compare compare, command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<KeyIterator>::value, "stable_sort_by_key needs device iterators");
    static_assert(is_device_iterator<ValueIterator>::value, "stable_sort_by_key needs device iterators");
    detail::sort_call(BCMI_STABLE_SORT_BY_KEY, keys_first, keys_last, values_first.device_ptr(), compare, queue);
}

template <class KeyIterator, class ValueIterator>
inline void stable_sort_by_key(KeyIterator keys_first, KeyIterator keys_last, ValueIterator values_first,
                               command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<KeyIterator>::value_type K;
    ::boost::compute::stable_sort_by_key(keys_first, keys_last, values_first, ::boost::compute::less<K>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
