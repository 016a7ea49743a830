//---------------------------------------------------------------------------//
// accumulate (reference: algorithm/accumulate.hpp:165-188): init op x0 op
// ... op x(n-1), returned to the host (blocking).  One parallel launch for
// every op; for float plus the reference folds serially on one work-item
// (serial_accumulate.hpp:38-40) while this build sums compensated with an
// FP64 combine (its error is reported against FP64).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_ACCUMULATE_HPP
#define BOOST_COMPUTE_ALGORITHM_ACCUMULATE_HPP

#include <boost/compute/algorithm/reduce.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class T, class BinaryFunction>
inline T accumulate(InputIterator first, InputIterator last, T init, BinaryFunction function,
                    command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<InputIterator>::value, "accumulate needs device iterators");
    typedef typename std::iterator_traits<InputIterator>::value_type V;
    static_assert(std::is_same<V, T>::value,
                  "accumulate: init must have the element type (a mixed-type fold is a serial fold)");
    if (first == last) return init;
    ::boost::compute::vector<T> tmp(1, queue.get_context());
    detail::bcmi_reduce_to(first, last, tmp.begin().device_ptr(), &init, function, queue);
    return tmp.begin().read(queue);
}

template <class InputIterator, class T>
inline T accumulate(InputIterator first, InputIterator last, T init, command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<InputIterator>::value_type V;
    return ::boost::compute::accumulate(first, last, init, ::boost::compute::plus<V>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
