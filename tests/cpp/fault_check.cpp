// Host-only check of the fault word -> status mapping a finished queue
// reports (compute_amd/csrc/fault.hpp) and of how the templates raise it:
// detail::bcmi_check turns BCMI_OUT_OF_RESOURCES into
// opencl_error(CL_OUT_OF_RESOURCES) carrying the message.  Built and run by
// tests/test_fault_report.py without a GPU.  Prints "<n> failures".
#include "../../compute_amd/csrc/fault.hpp"

#include <boost/compute/exception/opencl_error.hpp>

#include <cstdio>
#include <cstring>
#include <string>

static int failures = 0;
#define CHECK(c)                                                                    \
    do {                                                                            \
        if (!(c)) {                                                                 \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);                \
            ++failures;                                                             \
        }                                                                           \
    } while (0)

int main()
{
    using namespace bcmi;
    const char *msg = "unset";
    CHECK(fault_report(kFaultNone, &msg) == 0 && msg == nullptr);
    CHECK(fault_report(kFaultSpin, &msg) == BCMI_OUT_OF_RESOURCES && std::strstr(msg, "spin limit"));
    CHECK(fault_report(kFaultRange, &msg) == BCMI_OUT_OF_RESOURCES && std::strstr(msg, "outside the output range"));
    CHECK(fault_report(2, &msg) == BCMI_OUT_OF_RESOURCES && std::strstr(msg, "unknown fault"));
    CHECK(fault_report(0xFFFFFFFFu, &msg) == BCMI_OUT_OF_RESOURCES);
    // the status is the OpenCL code the reference raises
    CHECK(BCMI_OUT_OF_RESOURCES == -5);
    fault_report(kFaultSpin, &msg);
    const boost::compute::opencl_error e(BCMI_OUT_OF_RESOURCES, msg);
    CHECK(e.error_code() == -5);
    CHECK(std::string(e.what()).find("Out of Resources") == 0);
    CHECK(std::string(e.what()).find("spin limit") != std::string::npos);
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
