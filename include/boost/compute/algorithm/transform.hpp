//---------------------------------------------------------------------------//
// transform (reference: algorithm/transform.hpp:38-82): binary transforms
// with the recognised functors plus/minus/multiplies/divides/min/max, and
// the lambda a * _1 + _2 (saxpy, computed as fmaf(a, x, y): the contracted
// form of the reference's single OpenCL C expression "(a*x[i])+y[i]").
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_TRANSFORM_HPP
#define BOOST_COMPUTE_ALGORITHM_TRANSFORM_HPP

#include <boost/compute/functional/operator.hpp>
#include <boost/compute/lambda.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class T, class BinaryOperator>
inline void dispatch_transform(buffer_iterator<T> a, buffer_iterator<T> b, buffer_iterator<T> out, size_t n,
                               BinaryOperator, command_queue &queue)
{
    bcmi_check(bcmi_transform(queue.get(), bcmi_dtype<T>(), BinaryOperator::bcmi_op, a.device_ptr(), b.device_ptr(),
                              out.device_ptr(), n));
}

inline void dispatch_transform(buffer_iterator<float> x, buffer_iterator<float> y, buffer_iterator<float> out,
                               size_t n, lambda::saxpy_expr<float> e, command_queue &queue)
{
    bcmi_check(bcmi_saxpy(queue.get(), e.a, x.device_ptr(), y.device_ptr(), out.device_ptr(), n));
}

inline void dispatch_transform(buffer_iterator<float> x, buffer_iterator<float> y, buffer_iterator<float> out,
                               size_t n, lambda::saxpy_expr<double> e, command_queue &queue)
{
    // a double literal (2.5 * _1) is printed as a float expression by the
    // reference when the vectors are float: same kernel
    bcmi_check(bcmi_saxpy(queue.get(), (float)e.a, x.device_ptr(), y.device_ptr(), out.device_ptr(), n));
}

}  // namespace detail

template <class InputIterator1, class InputIterator2, class OutputIterator, class BinaryOperator>
inline OutputIterator transform(InputIterator1 first1, InputIterator1 last1, InputIterator2 first2,
                                OutputIterator result, BinaryOperator op,
                                command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator1>::value, "transform needs device iterators");
    static_assert(is_device_iterator<InputIterator2>::value, "transform needs device iterators");
    static_assert(is_device_iterator<OutputIterator>::value, "transform needs device iterators");
    const size_t n = (size_t)(last1 - first1);
    if (n) detail::dispatch_transform(first1, first2, result, n, op, queue);
    return result + (std::ptrdiff_t)n;
}

}  // namespace compute
}  // namespace boost

#endif
