//---------------------------------------------------------------------------//
// partition_copy (reference: algorithm/partition_copy.hpp): the elements for
// which predicate holds to first_true, the others to first_false, in input
// order, in one pass (the reference: two copy_if calls).  Blocking.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_PARTITION_COPY_HPP
#define BOOST_COMPUTE_ALGORITHM_PARTITION_COPY_HPP

#include <iterator>
#include <utility>

#include <boost/compute/detail/predicate.hpp>
#include <boost/compute/detail/select_count.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class OutputIterator1, class OutputIterator2, class UnaryPredicate>
inline std::pair<OutputIterator1, OutputIterator2>
partition_copy(InputIterator first, InputIterator last, OutputIterator1 first_true, OutputIterator2 first_false,
               UnaryPredicate predicate, command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "partition_copy needs device iterators");
    static_assert(is_device_iterator<OutputIterator1>::value, "partition_copy needs device iterators");
    static_assert(is_device_iterator<OutputIterator2>::value, "partition_copy needs device iterators");
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    const std::ptrdiff_t n = last - first;
    bcmi_predicate p = detail::make_predicate<T>(predicate);
    detail::select_count count(queue);
    detail::bcmi_check(bcmi_partition(queue.get(), detail::bcmi_dtype<T>(), first.device_ptr(),
                                      first_true.device_ptr(), first_false.device_ptr(), (size_t)n, &p, count.ptr()));
    const std::ptrdiff_t k = count.read(queue);
    return std::make_pair(first_true + k, first_false + (n - k));
}

}  // namespace compute
}  // namespace boost

#endif
