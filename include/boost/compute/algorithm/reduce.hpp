//---------------------------------------------------------------------------//
// reduce (reference: algorithm/reduce.hpp:275-305).  The result goes to a
// device iterator (enqueued) or to a host pointer (blocking read, as
// reduce.hpp:225).  Integers are exact; float plus is compensated FP32 with
// an FP64 cross-block combine (the reference's float association differs;
// its error is reported against FP64 instead).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_REDUCE_HPP
#define BOOST_COMPUTE_ALGORITHM_REDUCE_HPP

#include <boost/compute/container/vector.hpp>
#include <boost/compute/detail/materialize.hpp>
#include <boost/compute/functional/operator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T, class BinaryFunction>
inline void bcmi_reduce_to(buffer_iterator<T> first, buffer_iterator<T> last, T *out_device, const T *init,
                           BinaryFunction, command_queue &queue)
{
    bcmi_check(bcmi_reduce(queue.get(), bcmi_dtype<T>(), BinaryFunction::bcmi_op, first.device_ptr(),
                           (size_t)(last - first), init, out_device, 0, BCMI_PRECISION_COMPENSATED));
}

template <class T, class BinaryFunction>
inline void dispatch_reduce(buffer_iterator<T> first, buffer_iterator<T> last, buffer_iterator<T> result,
                            BinaryFunction f, command_queue &queue)
{
    bcmi_reduce_to(first, last, result.device_ptr(), (const T *)0, f, queue);
}

template <class T, class BinaryFunction>
inline void dispatch_reduce(buffer_iterator<T> first, buffer_iterator<T> last, T *result, BinaryFunction f,
                            command_queue &queue)
{
    ::boost::compute::vector<T> tmp(1, queue.get_context());
    bcmi_reduce_to(first, last, tmp.begin().device_ptr(), (const T *)0, f, queue);
    *result = tmp.begin().read(queue);
}

// counting / constant / transform iterator input: written out first
template <class InputIterator, class OutputIterator, class BinaryFunction>
inline void dispatch_reduce(InputIterator first, InputIterator last, OutputIterator result, BinaryFunction f,
                            command_queue &queue)
{
    const ::boost::compute::vector<typename iter_desc<InputIterator>::value_type> tmp =
        materialize(first, (size_t)(last - first), queue);
    dispatch_reduce(tmp.begin(), tmp.end(), result, f, queue);
    queue.finish();  // tmp is released on return
}

template <class InputIterator, class T, class BinaryFunction>
inline void bcmi_reduce_to(InputIterator first, InputIterator last, T *out_device, const T *init, BinaryFunction f,
                           command_queue &queue)
{
    const ::boost::compute::vector<T> tmp = materialize(first, (size_t)(last - first), queue);
    bcmi_reduce_to(tmp.begin(), tmp.end(), out_device, init, f, queue);
    queue.finish();
}

}  // namespace detail

template <class InputIterator, class OutputIterator, class BinaryFunction>
inline void reduce(InputIterator first, InputIterator last, OutputIterator result, BinaryFunction function,
                   command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "reduce needs device input iterators");
    if (first == last) return;
    detail::dispatch_reduce(first, last, result, function, queue);
}

template <class InputIterator, class OutputIterator>
inline void reduce(InputIterator first, InputIterator last, OutputIterator result,
                   command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    ::boost::compute::reduce(first, last, result, ::boost::compute::plus<T>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
