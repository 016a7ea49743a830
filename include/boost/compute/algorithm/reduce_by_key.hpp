//---------------------------------------------------------------------------//
// reduce_by_key (reference: algorithm/reduce_by_key.hpp:61-132): for each run
// of adjacent keys equal under equal_to, the run's last key and the fold of
// its values; returns the output ends (the run count is read back, blocking,
// as reduce_by_key_with_scan.hpp:485-487 does).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_REDUCE_BY_KEY_HPP
#define BOOST_COMPUTE_ALGORITHM_REDUCE_BY_KEY_HPP

#include <utility>

#include <boost/compute/container/vector.hpp>
#include <boost/compute/functional/operator.hpp>

namespace boost {
namespace compute {

template <class InputKeyIterator, class InputValueIterator, class OutputKeyIterator, class OutputValueIterator,
          class BinaryFunction, class BinaryPredicate>
inline std::pair<OutputKeyIterator, OutputValueIterator>
reduce_by_key(InputKeyIterator keys_first, InputKeyIterator keys_last, InputValueIterator values_first,
              OutputKeyIterator keys_result, OutputValueIterator values_result, BinaryFunction, BinaryPredicate,
              command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputKeyIterator>::value, "reduce_by_key needs device iterators");
    static_assert(is_device_iterator<InputValueIterator>::value, "reduce_by_key needs device iterators");
    static_assert(is_device_iterator<OutputKeyIterator>::value, "reduce_by_key needs device iterators");
    static_assert(is_device_iterator<OutputValueIterator>::value, "reduce_by_key needs device iterators");
    typedef typename std::iterator_traits<InputKeyIterator>::value_type K;
    typedef typename std::iterator_traits<InputValueIterator>::value_type V;
    static_assert(std::is_same<BinaryPredicate, equal_to<K> >::value,
                  "reduce_by_key: only equal_to<key_type> runs are supported");
    const size_t n = (size_t)(keys_last - keys_first);
    ::boost::compute::vector<unsigned long long> count(1, queue.get_context());
    detail::bcmi_check(bcmi_reduce_by_key(queue.get(), detail::bcmi_dtype<K>(), detail::bcmi_dtype<V>(),
                                          BinaryFunction::bcmi_op, keys_first.device_ptr(), values_first.device_ptr(),
                                          n, keys_result.device_ptr(), values_result.device_ptr(),
                                          (uint64_t *)count.begin().device_ptr(), BCMI_PRECISION_COMPENSATED));
    const std::ptrdiff_t r = (std::ptrdiff_t)count.begin().read(queue);
    return std::make_pair(keys_result + r, values_result + r);
}

template <class InputKeyIterator, class InputValueIterator, class OutputKeyIterator, class OutputValueIterator,
          class BinaryFunction>
inline std::pair<OutputKeyIterator, OutputValueIterator>
reduce_by_key(InputKeyIterator keys_first, InputKeyIterator keys_last, InputValueIterator values_first,
              OutputKeyIterator keys_result, OutputValueIterator values_result, BinaryFunction function,
              command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<InputKeyIterator>::value_type K;
    return ::boost::compute::reduce_by_key(keys_first, keys_last, values_first, keys_result, values_result, function,
                                           equal_to<K>(), queue);
}

template <class InputKeyIterator, class InputValueIterator, class OutputKeyIterator, class OutputValueIterator>
inline std::pair<OutputKeyIterator, OutputValueIterator>
reduce_by_key(InputKeyIterator keys_first, InputKeyIterator keys_last, InputValueIterator values_first,
              OutputKeyIterator keys_result, OutputValueIterator values_result,
              command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<InputKeyIterator>::value_type K;
    typedef typename std::iterator_traits<InputValueIterator>::value_type V;
    return ::boost::compute::reduce_by_key(keys_first, keys_last, values_first, keys_result, values_result,
                                           plus<V>(), equal_to<K>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
