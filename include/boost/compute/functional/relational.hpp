//---------------------------------------------------------------------------//
// signbit_ (reference: functional/relational.hpp:32, the OpenCL builtin
// signbit as a function object).  As a predicate it holds where the sign
// bit of a float is set (-0.0 and negative NaNs included).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_FUNCTIONAL_RELATIONAL_HPP
#define BOOST_COMPUTE_FUNCTIONAL_RELATIONAL_HPP

#include <cmath>

namespace boost {
namespace compute {

template <class T>
struct signbit_ {
    typedef int result_type;
    int operator()(const T &x) const { return std::signbit(x) ? 1 : 0; }
};

}  // namespace compute
}  // namespace boost

#endif
