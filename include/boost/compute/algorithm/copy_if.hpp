//---------------------------------------------------------------------------//
// copy_if and detail::copy_index_if (reference: algorithm/copy_if.hpp:25-62,
// through transform_if_impl, algorithm/transform_if.hpp:33-82): one
// single-pass compaction kernel instead of flags + scan + scatter.  The
// predicate is one of the forms detail/predicate.hpp recognises.  Blocking:
// the returned iterator needs the count.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_COPY_IF_HPP
#define BOOST_COMPUTE_ALGORITHM_COPY_IF_HPP

#include <iterator>
#include <type_traits>

#include <boost/compute/detail/predicate.hpp>
#include <boost/compute/detail/select_count.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class InputIterator, class OutputIterator, class Predicate>
inline OutputIterator select_into(InputIterator first, InputIterator last, OutputIterator result,
                                  const Predicate &predicate, int out_kind, command_queue &queue)
{
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    const size_t n = (size_t)(last - first);
    bcmi_predicate p = detail::make_predicate<T>(predicate);
    detail::select_count count(queue);
    detail::bcmi_check(bcmi_copy_if(queue.get(), detail::bcmi_dtype<T>(), first.device_ptr(), result.device_ptr(), n,
                                    &p, out_kind, count.ptr()));
    return result + count.read(queue);
}

// like copy_if() but writes the indices of the elements for which predicate
// holds (reference copy_if.hpp:25-41)
template <class InputIterator, class OutputIterator, class Predicate>
inline OutputIterator copy_index_if(InputIterator first, InputIterator last, OutputIterator result,
                                    Predicate predicate, command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "copy_index_if needs device iterators");
    static_assert(is_device_iterator<OutputIterator>::value, "copy_index_if needs device iterators");
    typedef typename std::iterator_traits<OutputIterator>::value_type I;
    static_assert(std::is_integral<I>::value && (sizeof(I) == 4 || sizeof(I) == 8),
                  "copy_index_if writes 32- or 64-bit integer indices");
    return select_into(first, last, result, predicate, sizeof(I) == 4 ? BCMI_SELECT_INDEX_U32 : BCMI_SELECT_INDEX_U64,
                       queue);
}

}  // namespace detail

template <class InputIterator, class OutputIterator, class Predicate>
inline OutputIterator copy_if(InputIterator first, InputIterator last, OutputIterator result, Predicate predicate,
                              command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "copy_if needs device iterators");
    static_assert(is_device_iterator<OutputIterator>::value, "copy_if needs device iterators");
    static_assert(std::is_same<typename std::iterator_traits<InputIterator>::value_type,
                               typename std::iterator_traits<OutputIterator>::value_type>::value,
                  "copy_if: input and output element types must match");
    return detail::select_into(first, last, result, predicate, BCMI_SELECT_VALUES, queue);
}

}  // namespace compute
}  // namespace boost

#endif
