// segscan_i64.hip -- reduce_by_key / segmented scans with int64_t keys.
#include "segscan.hpp"

namespace bcmi {
BCMI_SEG_DEFINE(seg_keys_i64, int64_t)
}  // namespace bcmi
