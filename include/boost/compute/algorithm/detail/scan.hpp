//---------------------------------------------------------------------------//
// detail::scan: the single-pass look-back scan behind exclusive_scan,
// inclusive_scan and partial_sum (reference: algorithm/detail/scan.hpp:22-39).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_DETAIL_SCAN_HPP
#define BOOST_COMPUTE_ALGORITHM_DETAIL_SCAN_HPP

#include <boost/compute/functional/operator.hpp>
#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {
namespace detail {

template <class InT, class OutT, class T, class BinaryOperator>
inline buffer_iterator<OutT> scan(buffer_iterator<InT> first, buffer_iterator<InT> last,
                                  buffer_iterator<OutT> result, bool exclusive, T init, BinaryOperator,
                                  command_queue &queue)
{
    static_assert(std::is_same<InT, OutT>::value, "scan: input and output element types must match");
    // minus and divides are not associative: a parallel scan has no
    // well-defined result for them (the reference's GPU scan folds in a tree
    // order, scan_on_gpu.hpp:84-93), so they are refused at compile time
    static_assert(BinaryOperator::bcmi_op == BCMI_PLUS || BinaryOperator::bcmi_op == BCMI_MULTIPLIES ||
                      BinaryOperator::bcmi_op == BCMI_MIN || BinaryOperator::bcmi_op == BCMI_MAX,
                  "scan: the operator must be plus, multiplies, min or max");
    const size_t n = (size_t)(last - first);
    const OutT ini = (OutT)init;
    bcmi_check(bcmi_scan(queue.get(), bcmi_dtype<OutT>(), BinaryOperator::bcmi_op, first.device_ptr(),
                         result.device_ptr(), n, exclusive ? &ini : 0, exclusive ? 1 : 0,
                         BCMI_PRECISION_COMPENSATED));
    return result + (std::ptrdiff_t)n;
}

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif
