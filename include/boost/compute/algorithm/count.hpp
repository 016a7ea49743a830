//---------------------------------------------------------------------------//
// count (reference: algorithm/count.hpp): count_if(_1 == value).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_COUNT_HPP
#define BOOST_COMPUTE_ALGORITHM_COUNT_HPP

#include <boost/compute/algorithm/count_if.hpp>
#include <boost/compute/lambda.hpp>

namespace boost {
namespace compute {

template <class InputIterator, class T>
inline size_t count(InputIterator first, InputIterator last, const T &value,
                    command_queue &queue = system::default_queue())
{
    return ::boost::compute::count_if(first, last, lambda::_1 == value, queue);
}

}  // namespace compute
}  // namespace boost

#endif
