//---------------------------------------------------------------------------//
// remove_if (reference: algorithm/remove_if.hpp): copy_if of not1(predicate)
// back into the range.  The reference copies the range to a temporary
// first; the compaction kernel works in place (a tile writes only below
// positions every earlier tile has read).  Blocking.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_REMOVE_IF_HPP
#define BOOST_COMPUTE_ALGORITHM_REMOVE_IF_HPP

#include <boost/compute/algorithm/copy_if.hpp>
#include <boost/compute/functional/logical.hpp>

namespace boost {
namespace compute {

template <class Iterator, class Predicate>
inline Iterator remove_if(Iterator first, Iterator last, Predicate predicate,
                          command_queue &queue = system::default_queue())
{
    static_assert(is_device_iterator<Iterator>::value, "remove_if needs device iterators");
    return detail::select_into(first, last, first, not1(predicate), BCMI_SELECT_VALUES, queue);
}

}  // namespace compute
}  // namespace boost

#endif
