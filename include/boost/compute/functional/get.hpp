//---------------------------------------------------------------------------//
// get<N> (reference: functional/get.hpp): component N of a vector-type
// element.  experimental::sort_by_transform reads it as the sort key.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_FUNCTIONAL_GET_HPP
#define BOOST_COMPUTE_FUNCTIONAL_GET_HPP

#include <cstddef>

namespace boost {
namespace compute {

template <std::size_t N>
struct get {
    template <class V>
    typename V::scalar_type operator()(const V &v) const
    {
        return v[N];
    }
};

}  // namespace compute
}  // namespace boost

#endif
