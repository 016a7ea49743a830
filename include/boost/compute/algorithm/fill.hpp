//---------------------------------------------------------------------------//
// fill (reference: algorithm/fill.hpp), by kernel so it is graph-safe.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_ALGORITHM_FILL_HPP
#define BOOST_COMPUTE_ALGORITHM_FILL_HPP

#include <boost/compute/system.hpp>
#include <boost/compute/type_traits/is_device_iterator.hpp>

namespace boost {
namespace compute {

template <class T>
inline void fill(buffer_iterator<T> first, buffer_iterator<T> last, const T &value,
                 command_queue &queue = system::default_queue())
{
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16,
                  "fill: element size must be 1, 2, 4, 8 or 16 bytes");
    const size_t n = (size_t)(last - first);
    if (n) detail::bcmi_check(bcmi_fill(queue.get(), first.device_ptr(), &value, sizeof(T), n));
}

template <class T, class Size>
inline void fill_n(buffer_iterator<T> first, Size count, const T &value,
                   command_queue &queue = system::default_queue())
{
    ::boost::compute::fill(first, first + (std::ptrdiff_t)count, value, queue);
}

}  // namespace compute
}  // namespace boost

#endif
