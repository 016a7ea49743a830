// segscan_u32.hip -- reduce_by_key / segmented scans with uint32_t keys.
#include "segscan.hpp"

namespace bcmi {
BCMI_SEG_DEFINE(seg_keys_u32, uint32_t)
}  // namespace bcmi
