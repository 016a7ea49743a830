// segscan_i32.hip -- reduce_by_key / segmented scans with int32_t keys.
#include "segscan.hpp"

namespace bcmi {
BCMI_SEG_DEFINE(seg_keys_i32, int32_t)
}  // namespace bcmi
