// segscan_f32.hip -- reduce_by_key / segmented scans with float keys.
#include "segscan.hpp"

namespace bcmi {
BCMI_SEG_DEFINE(seg_keys_f32, float)
}  // namespace bcmi
