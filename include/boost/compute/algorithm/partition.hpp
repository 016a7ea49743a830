//---------------------------------------------------------------------------//
// partition (reference: algorithm/partition.hpp): stable_partition.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_PARTITION_HPP
#define BOOST_COMPUTE_ALGORITHM_PARTITION_HPP

#include <boost/compute/algorithm/stable_partition.hpp>

namespace boost {
namespace compute {

template <class Iterator, class UnaryPredicate>
inline Iterator partition(Iterator first, Iterator last, UnaryPredicate predicate,
                          command_queue &queue = system::default_queue())
{
    return ::boost::compute::stable_partition(first, last, predicate, queue);
}

}  // namespace compute
}  // namespace boost

#endif
