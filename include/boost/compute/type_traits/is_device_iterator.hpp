//---------------------------------------------------------------------------//
// is_device_iterator: the compile-time gate of every algorithm (reference:
// iterator/buffer_iterator.hpp:273-275, e.g. sort_by_key.hpp:142-143).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_TYPE_TRAITS_IS_DEVICE_ITERATOR_HPP
#define BOOST_COMPUTE_TYPE_TRAITS_IS_DEVICE_ITERATOR_HPP

#include <type_traits>

#include <boost/compute/iterator/buffer_iterator.hpp>

namespace boost {
namespace compute {

template <class Iterator>
struct is_device_iterator : std::false_type {};

template <class T>
struct is_device_iterator<buffer_iterator<T> > : std::true_type {};

}  // namespace compute
}  // namespace boost

#endif
