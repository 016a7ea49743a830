//---------------------------------------------------------------------------//
// abs (reference: functional/integer.hpp:19, the OpenCL builtin abs as a
// function object).  In transform_reduce it maps to the |x| element map.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_FUNCTIONAL_INTEGER_HPP
#define BOOST_COMPUTE_FUNCTIONAL_INTEGER_HPP

namespace boost {
namespace compute {

template <class T>
struct abs {
    typedef T result_type;
    T operator()(const T &x) const { return x < T(0) ? T(-x) : x; }
};

}  // namespace compute
}  // namespace boost

#endif
