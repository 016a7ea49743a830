//---------------------------------------------------------------------------//
// unique_copy (reference: algorithm/unique_copy.hpp:99-175) with equal_to:
// the first element of each run of adjacent equal elements, in one
// compaction pass (the reference flags adjacent pairs, runs copy_index_if
// and gathers).  Blocking.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_UNIQUE_COPY_HPP
#define BOOST_COMPUTE_ALGORITHM_UNIQUE_COPY_HPP

#include <iterator>
#include <type_traits>

#include <boost/compute/detail/select_count.hpp>
#include <boost/compute/functional/operator.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class OutputIterator, class BinaryPredicate>
inline OutputIterator unique_copy(InputIterator first, InputIterator last, OutputIterator result, BinaryPredicate,
                                  command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "unique_copy needs device iterators");
    static_assert(is_device_iterator<OutputIterator>::value, "unique_copy needs device iterators");
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    static_assert(std::is_same<BinaryPredicate, equal_to<T> >::value, "unique_copy: only equal_to<T> is supported");
    const size_t n = (size_t)(last - first);
    detail::select_count count(queue);
    detail::bcmi_check(bcmi_unique(queue.get(), detail::bcmi_dtype<T>(), first.device_ptr(), result.device_ptr(), n,
                                   count.ptr()));
    return result + count.read(queue);
}

template <class InputIterator, class OutputIterator>
inline OutputIterator unique_copy(InputIterator first, InputIterator last, OutputIterator result,
                                  command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    return ::boost::compute::unique_copy(first, last, result, equal_to<T>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
