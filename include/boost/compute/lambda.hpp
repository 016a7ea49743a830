//---------------------------------------------------------------------------//
// The lambda placeholders _1, _2 (reference: lambda/placeholders.hpp).  The
// reference prints arbitrary lambda expressions as OpenCL C; this build has
// no run-time compiler, so it recognises the expression its benchmark uses,
// a * _1 + _2 (saxpy, perf/perf_saxpy.cpp:50-55), and maps it to a
// precompiled kernel.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_LAMBDA_HPP
#define BOOST_COMPUTE_LAMBDA_HPP

namespace boost {
namespace compute {
namespace lambda {

template <int I>
struct placeholder {
};

static const placeholder<1> _1 = placeholder<1>();
static const placeholder<2> _2 = placeholder<2>();

template <class T>
struct scaled_arg1 {
    T a;
};

template <class T>
struct saxpy_expr {
    T a;
};

template <class T>
inline scaled_arg1<T> operator*(T a, placeholder<1>)
{
    return scaled_arg1<T>{a};
}

template <class T>
inline saxpy_expr<T> operator+(scaled_arg1<T> s, placeholder<2>)
{
    return saxpy_expr<T>{s.a};
}

}  // namespace lambda
}  // namespace compute
}  // namespace boost

#endif
