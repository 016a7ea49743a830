//---------------------------------------------------------------------------//
// count_if (reference: algorithm/count_if.hpp, count_if_with_reduce.hpp):
// one streaming pass with per-block counts.  Blocking (returns size_t).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_COUNT_IF_HPP
#define BOOST_COMPUTE_ALGORITHM_COUNT_IF_HPP

#include <iterator>

#include <boost/compute/detail/predicate.hpp>
#include <boost/compute/detail/select_count.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class Predicate>
inline size_t count_if(InputIterator first, InputIterator last, Predicate predicate,
                       command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "count_if needs device iterators");
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    const size_t n = (size_t)(last - first);
    if (n == 0) return 0;
    bcmi_predicate p = detail::make_predicate<T>(predicate);
    detail::select_count count(queue);
    detail::bcmi_check(bcmi_count_if(queue.get(), detail::bcmi_dtype<T>(), first.device_ptr(), n, &p, count.ptr()));
    return (size_t)count.read(queue);
}

}  // namespace compute
}  // namespace boost

#endif
