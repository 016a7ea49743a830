//---------------------------------------------------------------------------//
// detail::scan: the single-pass look-back scan behind exclusive_scan,
// inclusive_scan and partial_sum (reference: algorithm/detail/scan.hpp:22-39).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_DETAIL_SCAN_HPP
#define BOOST_COMPUTE_ALGORITHM_DETAIL_SCAN_HPP

#include <boost/compute/container/vector.hpp>
#include <boost/compute/detail/iter_desc.hpp>
#include <boost/compute/functional/operator.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/types/fundamental.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class InT, class OutT, class T, class BinaryOperator>
inline buffer_iterator<OutT> scan(buffer_iterator<InT> first, buffer_iterator<InT> last,
                                  buffer_iterator<OutT> result, bool exclusive, T init, BinaryOperator,
                                  command_queue &queue)
{
    // minus and divides are not associative: a parallel scan has no
    // well-defined result for them (the reference's GPU scan folds in a tree
    // order, scan_on_gpu.hpp:84-93), so they are refused at compile time
    static_assert(BinaryOperator::bcmi_op == BCMI_PLUS || BinaryOperator::bcmi_op == BCMI_MULTIPLIES ||
                      BinaryOperator::bcmi_op == BCMI_MIN || BinaryOperator::bcmi_op == BCMI_MAX,
                  "scan: the operator must be plus, multiplies, min or max");
    const size_t n = (size_t)(last - first);
    // the scan computes in the input element type and converts on store
    // (scan_on_gpu.hpp:196-205, the local scan's scratch is of the input
    // type); init is taken in the input type as well
    const InT ini = (InT)init;
    if (std::is_same<InT, OutT>::value) {
        bcmi_check(bcmi_scan(queue.get(), bcmi_dtype<InT>(), BinaryOperator::bcmi_op, first.device_ptr(),
                             result.device_ptr(), n, exclusive ? &ini : 0, exclusive ? 1 : 0,
                             BCMI_PRECISION_COMPENSATED));
    } else if (n) {
        ::boost::compute::vector<InT> tmp(n, queue.get_context());
        bcmi_check(bcmi_scan(queue.get(), bcmi_dtype<InT>(), BinaryOperator::bcmi_op, first.device_ptr(),
                             tmp.begin().device_ptr(), n, exclusive ? &ini : 0, exclusive ? 1 : 0,
                             BCMI_PRECISION_COMPENSATED));
        bcmi_check(bcmi_convert(queue.get(), bcmi_dtype<InT>(), bcmi_dtype<OutT>(), tmp.begin().device_ptr(),
                                result.device_ptr(), n));
        queue.finish();  // tmp is released on return
    }
    return result + (std::ptrdiff_t)n;
}

template <class BinaryOperator>
inline void scan_op_check()
{
    static_assert(BinaryOperator::bcmi_op == BCMI_PLUS || BinaryOperator::bcmi_op == BCMI_MULTIPLIES ||
                      BinaryOperator::bcmi_op == BCMI_MIN || BinaryOperator::bcmi_op == BCMI_MAX,
                  "scan: the operator must be plus, multiplies, min or max");
}

// A counting_iterator, constant_iterator or transform_iterator input: the
// kernel generates or maps the elements as it reads them
// (test/test_scan.cpp:190-286).
template <class InputIterator, class OutT, class T, class BinaryOperator>
inline buffer_iterator<OutT> scan(InputIterator first, InputIterator last, buffer_iterator<OutT> result,
                                  bool exclusive, T init, BinaryOperator, command_queue &queue)
{
    typedef typename iter_desc<InputIterator>::value_type InT;
    scan_op_check<BinaryOperator>();
    const size_t n = (size_t)(last - first);
    const InT ini = (InT)init;
    const bcmi_iter d = iter_desc<InputIterator>::get(first);
    if (std::is_same<InT, OutT>::value) {
        bcmi_check(bcmi_scan_iter(queue.get(), bcmi_dtype<InT>(), 1, BinaryOperator::bcmi_op, &d,
                                  result.device_ptr(), n, exclusive ? &ini : 0, exclusive ? 1 : 0,
                                  BCMI_PRECISION_COMPENSATED));
    } else if (n) {
        // computed in the generated sequence's type, converted on store
        ::boost::compute::vector<InT> tmp(n, queue.get_context());
        bcmi_check(bcmi_scan_iter(queue.get(), bcmi_dtype<InT>(), 1, BinaryOperator::bcmi_op, &d,
                                  tmp.begin().device_ptr(), n, exclusive ? &ini : 0, exclusive ? 1 : 0,
                                  BCMI_PRECISION_COMPENSATED));
        bcmi_check(bcmi_convert(queue.get(), bcmi_dtype<InT>(), bcmi_dtype<OutT>(), tmp.begin().device_ptr(),
                                result.device_ptr(), n));
        queue.finish();  // tmp is released on return
    }
    return result + (std::ptrdiff_t)n;
}

// OpenCL vector elements (int2_, float4_ ...): plus, component by
// component (test/test_scan.cpp:140-180).
template <class S, std::size_t N, class T, class BinaryOperator>
inline buffer_iterator<vector_type<S, N> > scan(buffer_iterator<vector_type<S, N> > first,
                                                buffer_iterator<vector_type<S, N> > last,
                                                buffer_iterator<vector_type<S, N> > result, bool exclusive, T init,
                                                BinaryOperator, command_queue &queue)
{
    static_assert(N == 2 || N == 4, "scan: vector elements of 2 or 4 components");
    static_assert(BinaryOperator::bcmi_op == BCMI_PLUS, "scan: vector elements take plus");
    typedef vector_type<S, N> V;
    const size_t n = (size_t)(last - first);
    const V ini = V(init);
    bcmi_iter d;
    d.ptr = first.device_ptr();
    d.kind = BCMI_ITER_BUFFER;
    d.fn = BCMI_MAP_ID;
    d.value = 0;
    d.offset = 0;
    bcmi_check(bcmi_scan_iter(queue.get(), bcmi_dtype<S>(), (int)N, BCMI_PLUS, &d, result.device_ptr(), n,
                              exclusive ? &ini : 0, exclusive ? 1 : 0, BCMI_PRECISION_COMPENSATED));
    return result + (std::ptrdiff_t)n;
}

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif
