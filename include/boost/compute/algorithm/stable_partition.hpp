//---------------------------------------------------------------------------//
// stable_partition (reference: algorithm/stable_partition.hpp): the
// elements for which predicate holds first, the others after them, each
// group in input order; returns the end of the first group.  Two passes
// (the reference: a copy and two copy_if calls).  Blocking.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_STABLE_PARTITION_HPP
#define BOOST_COMPUTE_ALGORITHM_STABLE_PARTITION_HPP

#include <iterator>

#include <boost/compute/detail/predicate.hpp>
#include <boost/compute/detail/select_count.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class Iterator, class UnaryPredicate>
inline Iterator stable_partition(Iterator first, Iterator last, UnaryPredicate predicate,
                                 command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<Iterator>::value, "stable_partition needs device iterators");
    typedef typename std::iterator_traits<Iterator>::value_type T;
    const size_t n = (size_t)(last - first);
    bcmi_predicate p = detail::make_predicate<T>(predicate);
    detail::select_count count(queue);
    detail::bcmi_check(bcmi_partition(queue.get(), detail::bcmi_dtype<T>(), first.device_ptr(), first.device_ptr(),
                                      NULL, n, &p, count.ptr()));
    return first + count.read(queue);
}

}  // namespace compute
}  // namespace boost

#endif
