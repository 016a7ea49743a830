//---------------------------------------------------------------------------//
// transform (reference: algorithm/transform.hpp:38-82): binary transforms
// with the recognised functors plus/minus/multiplies/divides/min/max, and
// the lambda a * _1 + _2 (saxpy, computed as fmaf(a, x, y): the contracted
// form of the reference's single OpenCL C expression "(a*x[i])+y[i]").
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_TRANSFORM_HPP
#define BOOST_COMPUTE_ALGORITHM_TRANSFORM_HPP

#include <tuple>

#include <boost/compute/detail/iter_desc.hpp>
#include <boost/compute/functional/geometry.hpp>
#include <boost/compute/functional/get.hpp>
#include <boost/compute/functional/operator.hpp>
#include <boost/compute/iterator/zip_iterator.hpp>
#include <boost/compute/lambda.hpp>
#include <boost/compute/types/fundamental.hpp>
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

// binary transform with a counting / constant / transform iterator operand
template <class InputIterator1, class InputIterator2, class T, class BinaryOperator>
inline void dispatch_transform(InputIterator1 a, InputIterator2 b, buffer_iterator<T> out, size_t n, BinaryOperator,
                               command_queue &queue)
{
    const bcmi_iter da = iter_desc<InputIterator1>::get(a), db = iter_desc<InputIterator2>::get(b);
    bcmi_check(bcmi_transform_iter(queue.get(), bcmi_dtype<T>(), 1, BinaryOperator::bcmi_op, &da, &db,
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

// unary: out[i] = f(x[i]) as the transform_iterator of f over the input
template <class InputIterator, class T, class UnaryFunction>
inline void dispatch_unary(InputIterator first, size_t n, buffer_iterator<T> out, UnaryFunction f,
                           command_queue &queue)
{
    const bcmi_iter d = iter_desc<transform_iterator<InputIterator, UnaryFunction> >::get(
        ::boost::compute::make_transform_iterator(first, f));
    bcmi_check(bcmi_transform_iter(queue.get(), bcmi_dtype<T>(), 1, BCMI_PLUS, &d, 0, out.device_ptr(), n));
}

// get<N>() of a zip_iterator: component N's range, copied
template <class IteratorTuple, class T, std::size_t N>
inline void dispatch_unary(zip_iterator<IteratorTuple> first, size_t n, buffer_iterator<T> out, get<N>,
                           command_queue &queue)
{
    typedef typename std::tuple_element<N, IteratorTuple>::type It;
    const bcmi_iter d = iter_desc<It>::get(std::get<N>(first.get_iterator_tuple()));
    bcmi_check(bcmi_transform_iter(queue.get(), bcmi_dtype<T>(), 1, BCMI_PLUS, &d, 0, out.device_ptr(), n));
}

// length<floatN_>: the Euclidean length of each vector element
template <class S, std::size_t N, class T>
inline void dispatch_unary(buffer_iterator<vector_type<S, N> > first, size_t n, buffer_iterator<T> out,
                           length<vector_type<S, N> >, command_queue &queue)
{
    static_assert(std::is_same<S, float>::value && std::is_same<T, float>::value && (N == 2 || N == 4),
                  "length: float2_ or float4_ elements into floats");
    bcmi_iter d;
    d.ptr = first.device_ptr();
    d.kind = BCMI_ITER_BUFFER;
    d.fn = BCMI_MAP_LENGTH;
    d.value = 0;
    d.offset = 0;
    bcmi_check(bcmi_transform_iter(queue.get(), BCMI_F32, (int)N, BCMI_PLUS, &d, 0, out.device_ptr(), n));
}

}  // namespace detail

// transform(first, last, result, f): f one of the element maps a
// transform_iterator supports (_1 + c, abs<T>(), _1 * _1, ...), get<N>() of
// a zip_iterator, or length<floatN_>() (transform.hpp:38-53)
template <class InputIterator, class OutputIterator, class UnaryFunction>
inline OutputIterator transform(InputIterator first, InputIterator last, OutputIterator result, UnaryFunction f,
                                command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "transform needs device iterators");
    static_assert(is_device_iterator<OutputIterator>::value, "transform needs device iterators");
    const size_t n = (size_t)(last - first);
    if (n) detail::dispatch_unary(first, n, result, f, queue);
    return result + (std::ptrdiff_t)n;
}

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
