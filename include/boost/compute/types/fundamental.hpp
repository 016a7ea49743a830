//---------------------------------------------------------------------------//
// The OpenCL-named scalar and vector types (reference: types/fundamental.hpp).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_TYPES_FUNDAMENTAL_HPP
#define BOOST_COMPUTE_TYPES_FUNDAMENTAL_HPP

#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace boost {
namespace compute {

typedef int8_t char_;
typedef uint8_t uchar_;
typedef int16_t short_;
typedef uint16_t ushort_;
typedef int32_t int_;
typedef uint32_t uint_;
typedef int64_t long_;
typedef uint64_t ulong_;
typedef float float_;
typedef double double_;

// OpenCL vector types (reference: types/fundamental.hpp:53-180): N
// components, aligned to their size as in OpenCL, usable as vector
// elements (e.g. as the values of sort_by_key, or the records of
// experimental::sort_by_transform with get<N>).
template <class Scalar, std::size_t N>
class alignas(sizeof(Scalar) * N) vector_type {
public:
    typedef Scalar scalar_type;
    static const std::size_t size = N;

    vector_type()
    {
        for (std::size_t i = 0; i < N; ++i) m_s[i] = Scalar();
    }
    explicit vector_type(Scalar v)
    {
        for (std::size_t i = 0; i < N; ++i) m_s[i] = v;
    }
    template <class... Rest, class = typename std::enable_if<sizeof...(Rest) + 2 == N>::type>
    vector_type(Scalar s0, Scalar s1, Rest... rest) : m_s{s0, s1, static_cast<Scalar>(rest)...}
    {
    }
    Scalar &operator[](std::size_t i) { return m_s[i]; }
    const Scalar &operator[](std::size_t i) const { return m_s[i]; }
    bool operator==(const vector_type &o) const
    {
        for (std::size_t i = 0; i < N; ++i)
            if (!(m_s[i] == o.m_s[i])) return false;
        return true;
    }
    bool operator!=(const vector_type &o) const { return !(*this == o); }

private:
    Scalar m_s[N];
};

#define BOOST_COMPUTE_VECTOR_TYPEDEFS(S, NAME)      \
    typedef vector_type<S, 2> NAME##2_;            \
    typedef vector_type<S, 4> NAME##4_;            \
    typedef vector_type<S, 8> NAME##8_;            \
    typedef vector_type<S, 16> NAME##16_;
BOOST_COMPUTE_VECTOR_TYPEDEFS(char_, char)
BOOST_COMPUTE_VECTOR_TYPEDEFS(uchar_, uchar)
BOOST_COMPUTE_VECTOR_TYPEDEFS(short_, short)
BOOST_COMPUTE_VECTOR_TYPEDEFS(ushort_, ushort)
BOOST_COMPUTE_VECTOR_TYPEDEFS(int_, int)
BOOST_COMPUTE_VECTOR_TYPEDEFS(uint_, uint)
BOOST_COMPUTE_VECTOR_TYPEDEFS(long_, long)
BOOST_COMPUTE_VECTOR_TYPEDEFS(ulong_, ulong)
BOOST_COMPUTE_VECTOR_TYPEDEFS(float_, float)
BOOST_COMPUTE_VECTOR_TYPEDEFS(double_, double)
#undef BOOST_COMPUTE_VECTOR_TYPEDEFS

}  // namespace compute
}  // namespace boost

#endif
