// reduce_by_key.hip -- C-ABI entry points of reduce_by_key and the
// segmented (by-key) scans; the kernels are in segscan.hpp, instantiated
// per key type in segscan_<key>.hip.
#include "segscan.hpp"

namespace bcmi {

static int seg_dispatch(bcmi_queue q, int mode, int kdt, int vdt, int op, int precision, const void *keys,
                        const void *vals, void *out, void *keys_out, size_t n, const void *init, uint64_t *count_out)
{
    switch (kdt) {
    case BCMI_I32: return seg_keys_i32(q, mode, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_U32: return seg_keys_u32(q, mode, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_I64: return seg_keys_i64(q, mode, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_U64: return seg_keys_u64(q, mode, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_F32: return seg_keys_f32(q, mode, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out);
    case BCMI_F64: return seg_keys_f64(q, mode, vdt, op, precision, keys, vals, out, keys_out, n, init, count_out);
    default: set_error("reduce_by_key/segmented scan: key dtype %d unsupported", kdt); return BCMI_INVALID_VALUE;
    }
}

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_reduce_by_key(bcmi_queue q, int key_dtype, int value_dtype, int op, const void *keys,
                                  const void *values, size_t n, void *keys_out, void *values_out, uint64_t *count_out,
                                  int precision)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(count_out, BCMI_INVALID_VALUE, "reduce_by_key: null count");
    if (n == 0) {
        uint64_t zero = 0;
        return bcmi_memcpy(q, count_out, &zero, sizeof(zero), BCMI_H2D, 1);
    }
    BCMI_CHECK(keys && values && keys_out && values_out, BCMI_INVALID_VALUE, "reduce_by_key: null pointer");
    return seg_dispatch(q, SEG_RBK, key_dtype, value_dtype, op, precision, keys, values, values_out, keys_out, n,
                        nullptr, count_out);
}

extern "C" int bcmi_segmented_scan(bcmi_queue q, int key_dtype, int value_dtype, const void *keys,
                                   const void *values, void *out, size_t n, const void *init, int exclusive,
                                   int precision)
{
    BCMI_TRY(use_queue(q));
    if (n == 0) return 0;
    BCMI_CHECK(keys && values && out, BCMI_INVALID_VALUE, "segmented scan: null pointer");
    return seg_dispatch(q, exclusive ? SEG_EXCL : SEG_INCL, key_dtype, value_dtype, BCMI_PLUS, precision, keys,
                        values, out, nullptr, n, exclusive ? init : nullptr, nullptr);
}
