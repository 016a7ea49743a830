//---------------------------------------------------------------------------//
// unique (reference: algorithm/unique.hpp): unique_copy back into the
// range.  The reference copies the range to a temporary first; the kernel
// works in place.  Elements past the returned iterator keep their values.
// Blocking.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_UNIQUE_HPP
#define BOOST_COMPUTE_ALGORITHM_UNIQUE_HPP

#include <boost/compute/algorithm/unique_copy.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class BinaryPredicate>
inline InputIterator unique(InputIterator first, InputIterator last, BinaryPredicate op,
                            command_queue &queue = system::default_queue())
{
    return ::boost::compute::unique_copy(first, last, first, op, queue);
}

template <class InputIterator>
inline InputIterator unique(InputIterator first, InputIterator last, command_queue &queue = system::default_queue())
{
    typedef typename std::iterator_traits<InputIterator>::value_type T;
    return ::boost::compute::unique_copy(first, last, first, equal_to<T>(), queue);
}

}  // namespace compute
}  // namespace boost

#endif
