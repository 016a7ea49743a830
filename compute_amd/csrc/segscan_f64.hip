// segscan_f64.hip -- reduce_by_key / segmented scans with double keys.
#include "segscan.hpp"

namespace bcmi {
BCMI_SEG_DEFINE(seg_keys_f64, double)
}  // namespace bcmi
