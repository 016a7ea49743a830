//---------------------------------------------------------------------------//
// Header-only glue between the boost::compute:: templates and the C-ABI in
// bcmi.h: status -> exception mapping and C++ type -> bcmi_dtype mapping.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_DETAIL_BCMI_HPP
#define BOOST_COMPUTE_DETAIL_BCMI_HPP

#include <cstddef>
#include <cstdint>

#include <bcmi.h>

#include <boost/compute/exception/opencl_error.hpp>

namespace boost {
namespace compute {
namespace detail {

// Throws opencl_error carrying the status, as every failed enqueue in the
// reference does (command_queue.hpp:1494-1496).
inline void bcmi_check(int status)
{
    if (status != 0) {
        throw opencl_error(status, bcmi_last_error());
    }
}

template <class T>
struct bcmi_type {
    static const bool supported = false;
};

#define BOOST_COMPUTE_BCMI_TYPE(T, CODE)       \
    template <>                                \
    struct bcmi_type<T> {                      \
        static const bool supported = true;    \
        static int code() { return CODE; }     \
    };

BOOST_COMPUTE_BCMI_TYPE(char, BCMI_I8)  // char is signed on the supported hosts
BOOST_COMPUTE_BCMI_TYPE(signed char, BCMI_I8)
BOOST_COMPUTE_BCMI_TYPE(unsigned char, BCMI_U8)
BOOST_COMPUTE_BCMI_TYPE(short, BCMI_I16)
BOOST_COMPUTE_BCMI_TYPE(unsigned short, BCMI_U16)
BOOST_COMPUTE_BCMI_TYPE(int, BCMI_I32)
BOOST_COMPUTE_BCMI_TYPE(unsigned int, BCMI_U32)
BOOST_COMPUTE_BCMI_TYPE(long, BCMI_I64)
BOOST_COMPUTE_BCMI_TYPE(unsigned long, BCMI_U64)
BOOST_COMPUTE_BCMI_TYPE(long long, BCMI_I64)
BOOST_COMPUTE_BCMI_TYPE(unsigned long long, BCMI_U64)
BOOST_COMPUTE_BCMI_TYPE(float, BCMI_F32)
BOOST_COMPUTE_BCMI_TYPE(double, BCMI_F64)
#undef BOOST_COMPUTE_BCMI_TYPE

template <class T>
inline int bcmi_dtype()
{
    static_assert(bcmi_type<T>::supported, "element type not supported by the bcmi primitive core");
    return bcmi_type<T>::code();
}

}  // namespace detail
}  // namespace compute
}  // namespace boost

#endif
