// segscan_u64.hip -- reduce_by_key / segmented scans with uint64_t keys.
#include "segscan.hpp"

namespace bcmi {
BCMI_SEG_DEFINE(seg_keys_u64, uint64_t)
}  // namespace bcmi
