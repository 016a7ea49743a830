// fault.hpp -- what a finished queue reports for the word its kernels left
// in ctrl->fault.  Host-only (no HIP), so the mapping is compiled and tested
// on the CPU as well.
#pragma once

#include <cstdint>

#include "../../include/bcmi.h"

namespace bcmi {

// Codes the kernels store (relaxed, device scope) into ctrl->fault.
enum : uint32_t {
    kFaultNone = 0,
    kFaultSpin = 1,   // a look-back wait exceeded its spin limit and gave up
    kFaultRange = 3,  // an output range or index fell outside the output; the tile wrote nothing
};

// Status and message for a fault word.  Any non-zero word invalidates the
// results of the calls since the previous finish: BCMI_OUT_OF_RESOURCES,
// which the templates raise as opencl_error(CL_OUT_OF_RESOURCES).
inline int fault_report(uint32_t word, const char **msg)
{
    switch (word) {
    case kFaultNone: *msg = nullptr; return 0;
    case kFaultSpin: *msg = "a look-back wait exceeded its spin limit; results of the last calls are invalid"; break;
    case kFaultRange:
        *msg = "an output index fell outside the output range and was dropped; results are invalid";
        break;
    default: *msg = "a kernel reported an unknown fault; results of the last calls are invalid"; break;
    }
    return BCMI_OUT_OF_RESOURCES;
}

}  // namespace bcmi
