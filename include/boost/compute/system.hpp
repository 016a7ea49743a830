//---------------------------------------------------------------------------//
// system: default device / context / queue (reference: system.hpp:92-184).
// BOOST_COMPUTE_DEFAULT_DEVICE may name a device index or a substring of its
// name; with BOOST_COMPUTE_DEFAULT_ENFORCE=1 a missing device throws
// (system.hpp:244-274).  Initialised once, thread-safe (C++11 statics).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_SYSTEM_HPP
#define BOOST_COMPUTE_SYSTEM_HPP

#include <cstdlib>
#include <string>
#include <vector>

#include <boost/compute/command_queue.hpp>

namespace boost {
namespace compute {

class system
{
public:
    static size_t device_count()
    {
        int n = 0;
        detail::bcmi_check(bcmi_device_count(&n));
        return (size_t)n;
    }

    static std::vector<device> devices()
    {
        std::vector<device> out;
        for (size_t i = 0; i < device_count(); ++i) out.push_back(device((int)i));
        return out;
    }

    static device default_device()
    {
        static const device d = find_default_device();
        return d;
    }

    static context default_context()
    {
        static const context c(default_device());
        return c;
    }

    static command_queue &default_queue()
    {
        static command_queue q(default_context(), default_device());
        return q;
    }

    static void finish() { default_queue().finish(); }

private:
    static device find_default_device()
    {
        const size_t n = device_count();
        const char *want = std::getenv("BOOST_COMPUTE_DEFAULT_DEVICE");
        const char *enforce = std::getenv("BOOST_COMPUTE_DEFAULT_ENFORCE");
        if (want && *want) {
            char *end = 0;
            long idx = std::strtol(want, &end, 10);
            if (end && *end == 0 && idx >= 0 && (size_t)idx < n) return device((int)idx);
            for (size_t i = 0; i < n; ++i)
                if (device((int)i).name().find(want) != std::string::npos) return device((int)i);
            if (enforce && std::string(enforce) == "1") throw no_device_found();
        }
        if (n == 0) throw no_device_found();
        return device(0);
    }
};

}  // namespace compute
}  // namespace boost

#endif
