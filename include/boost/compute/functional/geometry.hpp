//---------------------------------------------------------------------------//
// length (reference: functional/geometry.hpp, the OpenCL builtin length of
// a float vector): the Euclidean length of a float2_ / float4_ element.
// transform_reduce and transform read it as the element map
// BCMI_MAP_LENGTH.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_FUNCTIONAL_GEOMETRY_HPP
#define BOOST_COMPUTE_FUNCTIONAL_GEOMETRY_HPP

#include <cmath>

namespace boost {
namespace compute {

template <class V>
struct length {
    typedef typename V::scalar_type result_type;
    result_type operator()(const V &v) const
    {
        result_type s = 0;
        for (std::size_t i = 0; i < V::size; ++i) s += v[i] * v[i];
        return std::sqrt(s);
    }
};

}  // namespace compute
}  // namespace boost

#endif
