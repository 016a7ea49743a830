//---------------------------------------------------------------------------//
// copy / copy_n between host ranges and device ranges (reference:
// algorithm/copy.hpp; host<->device copies are blocking, device->device
// copies are enqueued).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_COPY_HPP
#define BOOST_COMPUTE_ALGORITHM_COPY_HPP

#include <algorithm>
#include <iterator>
#include <vector>

#include <boost/compute/async/future.hpp>
#include <boost/compute/detail/iter_desc.hpp>
#include <boost/compute/event.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {
namespace detail {

// host -> device
template <class InputIterator, class T>
inline buffer_iterator<T> copy_dispatch(InputIterator first, InputIterator last, buffer_iterator<T> result,
                                        command_queue &queue, std::false_type, std::true_type)
{
    std::vector<T> tmp(first, last);
    if (!tmp.empty())
        bcmi_check(bcmi_memcpy(queue.get(), result.device_ptr(), tmp.data(), tmp.size() * sizeof(T), BCMI_H2D, 1));
    return result + (std::ptrdiff_t)tmp.size();
}

// device -> host
template <class T, class OutputIterator>
inline OutputIterator copy_dispatch(buffer_iterator<T> first, buffer_iterator<T> last, OutputIterator result,
                                    command_queue &queue, std::true_type, std::false_type)
{
    const size_t n = (size_t)(last - first);
    std::vector<T> tmp(n);
    if (n) bcmi_check(bcmi_memcpy(queue.get(), tmp.data(), first.device_ptr(), n * sizeof(T), BCMI_D2H, 1));
    return std::copy(tmp.begin(), tmp.end(), result);
}

// device -> device
template <class T>
inline buffer_iterator<T> copy_dispatch(buffer_iterator<T> first, buffer_iterator<T> last, buffer_iterator<T> result,
                                        command_queue &queue, std::true_type, std::true_type)
{
    const size_t n = (size_t)(last - first);
    if (n)
        bcmi_check(bcmi_memcpy(queue.get(), result.device_ptr(), first.device_ptr(), n * sizeof(T), BCMI_D2D, 0));
    return result + (std::ptrdiff_t)n;
}

// device -> device between element types: converted element by element on
// the device (copy_on_device.hpp:59-92 assigns through the element types)
template <class S, class D>
inline typename std::enable_if<!std::is_same<S, D>::value, buffer_iterator<D> >::type copy_dispatch(
    buffer_iterator<S> first, buffer_iterator<S> last, buffer_iterator<D> result, command_queue &queue,
    std::true_type, std::true_type)
{
    const size_t n = (size_t)(last - first);
    if (n)
        bcmi_check(bcmi_convert(queue.get(), bcmi_dtype<S>(), bcmi_dtype<D>(), first.device_ptr(),
                                result.device_ptr(), n));
    return result + (std::ptrdiff_t)n;
}

// counting / constant / transform iterator into a device range: the
// sequence generated (and mapped) on the device (test/test_constant_iterator.cpp:60-91)
template <class InputIterator, class T>
inline buffer_iterator<T> copy_dispatch(InputIterator first, InputIterator last, buffer_iterator<T> result,
                                        command_queue &queue, std::true_type, std::true_type)
{
    static_assert(std::is_same<typename iter_desc<InputIterator>::value_type, T>::value,
                  "copy: a generated range is copied into elements of its own type");
    const size_t n = (size_t)(last - first);
    const bcmi_iter d = iter_desc<InputIterator>::get(first);
    if (n) bcmi_check(bcmi_transform_iter(queue.get(), bcmi_dtype<T>(), 1, BCMI_PLUS, &d, 0, result.device_ptr(), n));
    return result + (std::ptrdiff_t)n;
}

}  // namespace detail

template <class InputIterator, class OutputIterator>
inline OutputIterator copy(InputIterator first, InputIterator last, OutputIterator result,
                           command_queue &queue = system::default_queue())
{
    return detail::copy_dispatch(first, last, result, queue,
                                 std::integral_constant<bool, is_device_iterator<InputIterator>::value>(),
                                 std::integral_constant<bool, is_device_iterator<OutputIterator>::value>());
}

// copy_async (reference: algorithm/copy.hpp:102-150 dispatch_copy_async):
// an asynchronous copy between a contiguous host range (a pointer) and a
// device range, or between device ranges, on the queue's stream.  The host
// memory must stay valid until the future is waited on; page-locked host
// memory (bcmi_host_malloc) lets the transfer overlap host work.
template <class T>
inline future<buffer_iterator<T> > copy_async(const T *first, const T *last, buffer_iterator<T> result,
                                              command_queue &queue = system::default_queue())
{
    const size_t n = (size_t)(last - first);
    if (n) detail::bcmi_check(bcmi_memcpy(queue.get(), result.device_ptr(), first, n * sizeof(T), BCMI_H2D, 0));
    return future<buffer_iterator<T> >(result + (std::ptrdiff_t)n, event::marker(queue));
}

template <class T>
inline future<T *> copy_async(buffer_iterator<T> first, buffer_iterator<T> last, T *result,
                              command_queue &queue = system::default_queue())
{
    const size_t n = (size_t)(last - first);
    if (n) detail::bcmi_check(bcmi_memcpy(queue.get(), result, first.device_ptr(), n * sizeof(T), BCMI_D2H, 0));
    return future<T *>(result + n, event::marker(queue));
}

template <class T>
inline future<buffer_iterator<T> > copy_async(buffer_iterator<T> first, buffer_iterator<T> last,
                                              buffer_iterator<T> result,
                                              command_queue &queue = system::default_queue())
{
    const size_t n = (size_t)(last - first);
    if (n)
        detail::bcmi_check(
            bcmi_memcpy(queue.get(), result.device_ptr(), first.device_ptr(), n * sizeof(T), BCMI_D2D, 0));
    return future<buffer_iterator<T> >(result + (std::ptrdiff_t)n, event::marker(queue));
}

template <class InputIterator, class Size, class OutputIterator>
inline OutputIterator copy_n(InputIterator first, Size count, OutputIterator result,
                             command_queue &queue = system::default_queue())
{
    InputIterator last = first;
    std::advance(last, count);
    return ::boost::compute::copy(first, last, result, queue);
}

}  // namespace compute
}  // namespace boost

#endif
