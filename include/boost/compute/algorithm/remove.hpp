//---------------------------------------------------------------------------//
// remove (reference: algorithm/remove.hpp): remove_if(_1 == value).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_REMOVE_HPP
#define BOOST_COMPUTE_ALGORITHM_REMOVE_HPP

#include <boost/compute/algorithm/remove_if.hpp>
#include <boost/compute/lambda.hpp>

namespace boost {
namespace compute {

template <class Iterator, class T>
inline Iterator remove(Iterator first, Iterator last, const T &value, command_queue &queue = system::default_queue())
{
    return ::boost::compute::remove_if(first, last, lambda::_1 == value, queue);
}

}  // namespace compute
}  // namespace boost

#endif
