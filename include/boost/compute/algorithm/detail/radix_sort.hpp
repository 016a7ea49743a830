//---------------------------------------------------------------------------//
// detail::radix_sort / radix_sort_by_key (reference:
// algorithm/detail/radix_sort.hpp:428-462): always the radix path, at any n,
// as the reference's own tests call them (test/test_radix_sort.cpp:46).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_DETAIL_RADIX_SORT_HPP
#define BOOST_COMPUTE_ALGORITHM_DETAIL_RADIX_SORT_HPP

#include <boost/compute/functional/operator.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class K, class V>
inline void bcmi_sort_call(int api, buffer_iterator<K> first, buffer_iterator<K> last, V *values, bool descending,
                           command_queue &queue)
{
    const size_t n = (size_t)(last - first);
    bcmi_check(bcmi_sort(queue.get(), api, bcmi_dtype<K>(), first.device_ptr(), values, values ? sizeof(V) : 0, n,
                         descending ? 1 : 0));
}

template <class T>
inline void radix_sort(buffer_iterator<T> first, buffer_iterator<T> last, command_queue &queue)
{
    bcmi_sort_call<T, T>(BCMI_RADIX_SORT, first, last, (T *)0, false, queue);
}

template <class T>
inline void radix_sort(buffer_iterator<T> first, buffer_iterator<T> last, const bool ascending, command_queue &queue)
{
    bcmi_sort_call<T, T>(BCMI_RADIX_SORT, first, last, (T *)0, !ascending, queue);
}

template <class K, class V>
inline void radix_sort_by_key(buffer_iterator<K> keys_first, buffer_iterator<K> keys_last,
                              buffer_iterator<V> values_first, command_queue &queue)
{
    bcmi_sort_call<K, V>(BCMI_RADIX_SORT, keys_first, keys_last, values_first.device_ptr(), false, queue);
}

template <class K, class V>
inline void radix_sort_by_key(buffer_iterator<K> keys_first, buffer_iterator<K> keys_last,
                              buffer_iterator<V> values_first, const bool ascending, command_queue &queue)
{
    bcmi_sort_call<K, V>(BCMI_RADIX_SORT, keys_first, keys_last, values_first.device_ptr(), !ascending, queue);
}

// less<T> -> ascending, greater<T> -> descending; comparator expressions take
// the comparison sort instead (merge_sort_on_gpu.hpp).
template <class C, class T>
struct sort_direction;
template <class T>
struct sort_direction<less<T>, T> {
    static const bool descending = false;
};
template <class T>
struct sort_direction<greater<T>, T> {
    static const bool descending = true;
};

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif
