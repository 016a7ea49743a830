// segmented.hip -- batched segmented sort (filled in below).
#include "device.hpp"

using namespace bcmi;

extern "C" int bcmi_segmented_sort(bcmi_queue q, int key_dtype, void *keys, void *values, size_t value_bytes,
                                   const uint64_t *offsets, const uint64_t *host_offsets, size_t nseg, int descending)
{
    BCMI_TRY(use_queue(q));
    set_error("segmented sort: not available yet");
    return BCMI_INVALID_OPERATION;
}
