//---------------------------------------------------------------------------//
// The OpenCL-named scalar typedefs (reference: types/fundamental.hpp).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_TYPES_FUNDAMENTAL_HPP
#define BOOST_COMPUTE_TYPES_FUNDAMENTAL_HPP

#include <cstdint>

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

}  // namespace compute
}  // namespace boost

#endif
