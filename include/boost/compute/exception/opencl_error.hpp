//---------------------------------------------------------------------------//
// opencl_error: the exception every failed call throws.  Same interface as
// the reference's (exception/opencl_error.hpp:30-62): error_code() is the
// (OpenCL-valued) status, what() the message.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_EXCEPTION_OPENCL_ERROR_HPP
#define BOOST_COMPUTE_EXCEPTION_OPENCL_ERROR_HPP

#include <exception>
#include <sstream>
#include <string>

namespace boost {
namespace compute {

class opencl_error : public std::exception
{
public:
    explicit opencl_error(int error, const char *detail = 0) : m_error(error)
    {
        std::ostringstream s;
        s << to_string(error);
        if (detail && *detail) s << ": " << detail;
        m_error_string = s.str();
    }

    ~opencl_error() throw() {}

    int error_code() const throw() { return m_error; }
    std::string error_string() const throw() { return m_error_string; }
    const char *what() const throw() { return m_error_string.c_str(); }

    static std::string to_string(int error)
    {
        switch (error) {
        case 0: return "Success";
        case -1: return "Device Not Found";
        case -4: return "Memory Object Allocation Failure";
        case -5: return "Out of Resources";
        case -30: return "Invalid Value";
        case -36: return "Invalid Command Queue";
        case -59: return "Invalid Operation";
        case -61: return "Invalid Buffer Size";
        default: {
            std::ostringstream s;
            s << "Unknown OpenCL Error (" << error << ")";
            return s.str();
        }
        }
    }

private:
    int m_error;
    std::string m_error_string;
};

class no_device_found : public std::exception
{
public:
    const char *what() const throw() { return "No OpenCL device found"; }
};

}  // namespace compute
}  // namespace boost

#endif
