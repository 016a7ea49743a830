/*
 * bcmi.h -- C-ABI of the MI355X-native data-parallel primitive core.
 *
 * The reference (Boost.Compute 0.5.0) has no C-ABI: its boundary is the
 * header-only C++ templates in namespace boost::compute over vector
 * iterators and a command_queue.  This library is the native layer those
 * templates call in this build; include/boost/compute/ holds the
 * source-compatible templates (the "reference-side binding").  Every entry
 * point below names the reference interface it replaces.
 *
 * Conventions
 *   - Plain pointers and sizes; no C++ or torch types.
 *   - Every function returns an int status: 0 on success, otherwise a
 *     negative code whose value matches the OpenCL error the reference would
 *     raise (exception/opencl_error.hpp:30-62); bcmi_last_error() gives a
 *     thread-local message.
 *   - Device pointers come from bcmi_malloc (hipMalloc).  A
 *     buffer_iterator's (buffer, index) pair is passed as one element
 *     pointer, so no offset is ever baked into code.
 *   - Algorithms are asynchronous on the queue (one HIP stream) unless the
 *     comment says "blocking", mirroring which reference calls block
 *     (reduce.hpp:225, accumulate.hpp:53-55,
 *     reduce_by_key_with_scan.hpp:485-487).
 *   - Sizes are size_t; sorts accept n <= 2^32 - 2^20 (32-bit tile and
 *     digit counts, as radix_sort.hpp:131-137 uses uint counts).
 */
#ifndef BCMI_H
#define BCMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
enum bcmi_status {
    BCMI_SUCCESS = 0,
    BCMI_DEVICE_NOT_FOUND = -1,              /* CL_DEVICE_NOT_FOUND */
    BCMI_MEM_ALLOCATION_FAILURE = -4,        /* CL_MEM_OBJECT_ALLOCATION_FAILURE */
    BCMI_OUT_OF_RESOURCES = -5,              /* CL_OUT_OF_RESOURCES */
    BCMI_INVALID_VALUE = -30,                /* CL_INVALID_VALUE */
    BCMI_INVALID_COMMAND_QUEUE = -36,        /* CL_INVALID_COMMAND_QUEUE */
    BCMI_INVALID_OPERATION = -59,            /* CL_INVALID_OPERATION */
    BCMI_INVALID_BUFFER_SIZE = -61,          /* CL_INVALID_BUFFER_SIZE */
    BCMI_RUNTIME_ERROR = -9999               /* HIP runtime failure */
};

/* element types (a key or value type is named by one of these) */
enum bcmi_dtype {
    BCMI_I8 = 1, BCMI_U8, BCMI_I16, BCMI_U16, BCMI_I32, BCMI_U32,
    BCMI_I64, BCMI_U64, BCMI_F32, BCMI_F64
};

/* binary functors recognised by the dispatch: plus<T>, multiplies<T>,
 * min<T>, max<T>, minus<T>, divides<T> (functional/operator.hpp:57-96,
 * functional/common.hpp) */
enum bcmi_op {
    BCMI_PLUS = 0, BCMI_MULTIPLIES = 1, BCMI_MIN = 2, BCMI_MAX = 3,
    BCMI_MINUS = 4, BCMI_DIVIDES = 5
};

/* sort front-ends (which reference template is being served) */
enum bcmi_sort_api {
    BCMI_SORT = 0,               /* sort.hpp:182-202: n<=32 insertion sort, else radix */
    BCMI_SORT_BY_KEY = 1,        /* sort_by_key.hpp:135-163: n<32 insertion, else radix */
    BCMI_STABLE_SORT = 2,        /* stable_sort.hpp:81-110: radix at any n */
    BCMI_STABLE_SORT_BY_KEY = 3, /* stable_sort_by_key.hpp:135-163: n<32 insertion */
    BCMI_RADIX_SORT = 4          /* detail::radix_sort(_by_key), radix_sort.hpp:428-462 */
};

/* float reduction/scan accumulation */
enum bcmi_precision {
    BCMI_PRECISION_COMPENSATED = 0, /* two-sum FP32 in waves, FP64 cross-block (default) */
    BCMI_PRECISION_NAIVE = 1        /* plain FP32 tree (for error comparison only) */
};

/* ---------------------------------------------------- errors / devices */
const char *bcmi_last_error(void);
int bcmi_get_version(int *major, int *minor);

typedef struct bcmi_device_info {
    char name[256];          /* marketing name */
    char arch[64];           /* gcnArchName, e.g. gfx950 */
    int compute_units;       /* CL_DEVICE_MAX_COMPUTE_UNITS analogue */
    int wavefront_size;
    int clock_khz;
    size_t global_mem_bytes;
    size_t lds_bytes_per_block;
    int pci_bus_id;
} bcmi_device_info;

/* device.hpp / system.hpp:92-184 (default device = device 0 unless
 * BOOST_COMPUTE_DEFAULT_DEVICE names another, system.hpp:244-274) */
int bcmi_device_count(int *count);
int bcmi_device_get_info(int device, bcmi_device_info *info);

/* -------------------------------------------------------------- queues */
/* command_queue.hpp: an in-order queue = one HIP stream plus the queue's
 * scratch arena (temporaries) and look-back state.  Not thread-safe: like a
 * command_queue, use one per host thread. */
typedef struct bcmi_queue_s *bcmi_queue;
int bcmi_queue_create(int device, bcmi_queue *queue);
int bcmi_queue_destroy(bcmi_queue queue);
/* command_queue::finish (command_queue.hpp:1564): blocking.  Returns
 * BCMI_OUT_OF_RESOURCES when a kernel since the previous finish gave up a
 * bounded look-back wait or dropped an out-of-range output (the results of
 * those calls are invalid; the fault is cleared). */
int bcmi_queue_finish(bcmi_queue queue);
/* Test support, no counterpart in the reference: runs the bounded wait the
 * look-backs use on a word nothing sets, with the queue's spin limit lowered
 * to spin_limit polls (1..65536) for that launch, so the next
 * bcmi_queue_finish reports the give-up path.  No kernel faults or hangs. */
int bcmi_debug_fault_probe(bcmi_queue queue, unsigned spin_limit);
int bcmi_queue_device(bcmi_queue queue, int *device);
/* the underlying hipStream_t, for interop */
void *bcmi_queue_native(bcmi_queue queue);
/* Pre-size the scratch arena and look-back state for n elements of up to
 * elem_bytes key + value bytes (call before hipGraph capture). */
int bcmi_queue_reserve(bcmi_queue queue, size_t n, size_t elem_bytes);

/* ------------------------------------------------------------- memory */
/* buffer.hpp:53-121 (clCreateBuffer / release) */
int bcmi_malloc(size_t bytes, void **ptr);
int bcmi_free(void *ptr);
/* command_queue::enqueue_{read,write,copy}_buffer (command_queue.hpp:297-766);
 * blocking != 0 waits for completion */
enum bcmi_copy_kind { BCMI_H2D = 1, BCMI_D2H = 2, BCMI_D2D = 3 };
int bcmi_memcpy(bcmi_queue queue, void *dst, const void *src, size_t bytes,
                int kind, int blocking);
/* fill (algorithm/fill.hpp): pattern of 1, 2, 4, 8 or 16 bytes, by kernel
 * (safe under graph capture) */
int bcmi_fill(bcmi_queue queue, void *dst, const void *pattern,
              size_t pattern_bytes, size_t count);

/* Pinned (page-locked) host memory, the analogue of the reference's
 * host-pointer buffers (buffer.hpp CL_MEM_ALLOC_HOST_PTR,
 * allocator/pinned_allocator.hpp): bcmi_memcpy with blocking == 0 between it
 * and device memory runs asynchronously on the queue's stream (copy_async,
 * algorithm/copy.hpp:102-150), at full link rate. */
int bcmi_host_malloc(size_t bytes, void **ptr);
int bcmi_host_free(void *ptr);

/* --------------------------------------------------------------- events */
/* event.hpp:86-89,204,264 (profiling duration) */
typedef struct bcmi_event_s *bcmi_event;
int bcmi_event_create(bcmi_event *event);
int bcmi_event_destroy(bcmi_event event);
int bcmi_event_record(bcmi_event event, bcmi_queue queue);
int bcmi_event_synchronize(bcmi_event event);
int bcmi_event_elapsed_ms(bcmi_event start, bcmi_event end, float *ms);

/* ----------------------------------------------------------- algorithms */
/* sort / sort_by_key / stable_sort / stable_sort_by_key / detail::radix_sort.
 * keys: n elements of key_dtype, sorted in place.  values: NULL, or n
 * elements of value_bytes (1, 2, 4, 8 or 16) permuted with the keys.
 * descending != 0 serves greater<T>.  The result is bit-identical to the
 * reference's GPU column: radix order by the radix() key transform
 * (radix_sort.hpp:98-129; -0.0 before +0.0, NaNs by sign at the ends) and,
 * below the api's threshold, the one-work-item insertion sort
 * (insertion_sort.hpp:25-159). */
int bcmi_sort(bcmi_queue queue, int api, int key_dtype, void *keys,
              void *values, size_t value_bytes, size_t n, int descending);

/* bcmi_sort with no allocation on any path.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, order and results as bcmi_sort; every
 * temporary (key and value copies, histograms, look-back words) lives in a
 * 64 MiB device array of the library, one copy per GPU, held until the
 * process ends.  A sort whose temporaries exceed it (about 15 M 4-byte keys,
 * 7 M 4-byte pairs) fails with BCMI_INVALID_BUFFER_SIZE before any work is
 * enqueued.  All calls on one GPU share the array: calls that overlap on
 * that GPU (two queues at once) race.  The queue's kernel profiler does not
 * record these calls. */
int bcmi_sort_noalloc(bcmi_queue queue, int api, int key_dtype, void *keys,
                      void *values, size_t value_bytes, size_t n, int descending);

/* exclusive_scan (exclusive_scan.hpp:55-104), inclusive_scan
 * (inclusive_scan.hpp:53-87), partial_sum (partial_sum.hpp:31-41).
 * in == out is allowed.  init: host pointer to one element (exclusive only;
 * NULL means the op's identity for inclusive).  op is BCMI_PLUS,
 * BCMI_MULTIPLIES, BCMI_MIN or BCMI_MAX (MINUS and DIVIDES are not
 * associative and give BCMI_INVALID_VALUE; the templates refuse them at
 * compile time).  Integers wrap; float plus is compensated (precision)
 * with an FP64 carry between tiles. */
int bcmi_scan(bcmi_queue queue, int dtype, int op, const void *in, void *out,
              size_t n, const void *init, int exclusive, int precision);

/* reduce (reduce.hpp:275-305) and accumulate (accumulate.hpp:165-188).
 * Writes op-fold of [in, in+n) (then folded with *init when init != NULL,
 * i.e. accumulate) to the device pointer out (one element).  Float plus is
 * compensated with an FP64 cross-block combine; out_f64 (may be NULL)
 * receives the FP64 value before the final rounding. */
int bcmi_reduce(bcmi_queue queue, int dtype, int op, const void *in, size_t n,
                const void *init, void *out, double *out_f64, int precision);

/* reduce_by_key (reduce_by_key.hpp:61-132) with equal_to on keys: for each
 * run of adjacent equal keys, the run's last key and the op-fold of its
 * values.  *count_out (device pointer, uint64) receives the number of runs;
 * the host-side shim reads it back (blocking, as the reference does). */
int bcmi_reduce_by_key(bcmi_queue queue, int key_dtype, int value_dtype,
                       int op, const void *keys, const void *values, size_t n,
                       void *keys_out, void *values_out, uint64_t *count_out,
                       int precision);

/* transform (transform.hpp:38-82) with a recognised binary functor:
 * out[i] = op(a[i], b[i]). */
int bcmi_transform(bcmi_queue queue, int dtype, int op, const void *a,
                   const void *b, void *out, size_t n);
/* transform with the lambda a*_1+_2 (perf/perf_saxpy.cpp:50-55):
 * out[i] = fmaf(a, x[i], y[i]) (the contracted form of "(a*x[i])+y[i]"). */
int bcmi_saxpy(bcmi_queue queue, float a, const float *x, const float *y,
               float *out, size_t n);

/* ------------------------------------------------- batched / segmented */
/* Batched segmented sort.  Segment s is [host_offsets[s], host_offsets[s+1])
 * of keys/values (value_bytes: 0, 1, 2, 4, 8 or 16); every segment ends up
 * exactly as one stable_sort_by_key call on it would leave it
 * (stable_sort_by_key.hpp:29-77: below 32 elements the insertion sort, else
 * radix order).  No reference counterpart as one call: it replaces one
 * stable_sort_by_key call per segment.  bcmi_segmented_sort plans, runs and
 * waits (blocking).  For repeated or graph-captured use, build the plan once:
 * running a plan does no host work and allocates nothing. */
typedef struct bcmi_segplan_s *bcmi_segplan;
int bcmi_segmented_sort(bcmi_queue queue, int key_dtype, void *keys,
                        void *values, size_t value_bytes,
                        const uint64_t *host_offsets, size_t nseg,
                        int descending);
int bcmi_segmented_plan_create(bcmi_queue queue, int key_dtype,
                               size_t value_bytes, const uint64_t *host_offsets,
                               size_t nseg, int descending, bcmi_segplan *plan);
int bcmi_segmented_sort_plan(bcmi_queue queue, bcmi_segplan plan, void *keys,
                             void *values);
int bcmi_segmented_plan_destroy(bcmi_segplan plan);

/* Segmented scan over runs of adjacent equal keys (the structure of
 * reduce_by_key_with_scan.hpp:104-293 exposed as a scan): exclusive runs
 * start at *init, inclusive runs restart at each run head. */
int bcmi_segmented_scan(bcmi_queue queue, int key_dtype, int value_dtype,
                        const void *keys, const void *values, void *out,
                        size_t n, const void *init, int exclusive,
                        int precision);

/* ----------------------------------------------- selection (compaction) */
/* The unary predicates the dispatch recognises.  The reference prints a
 * lambda expression as OpenCL C (lambda/context.hpp:145-156); this build
 * has no run-time compiler and recognises the forms
 *     f(_1) cmp c        with f = _1, _1 * a, _1 % a, _1 & a or signbit(_1)
 * (signbit_: functional/relational.hpp:32), cmp one of < <= > >= == !=,
 * optionally negated (not1: functional/logical.hpp:113-175).  a and c hold
 * one value of the comparison type in their low bytes: int32 for 8- and
 * 16-bit integers (C's integer promotion) and for signbit, else the element
 * type.  % and & need an integer type, signbit a float type; % by 0 is
 * rejected. */
enum bcmi_pred_fn {
    BCMI_FN_ID = 0, BCMI_FN_MUL = 1, BCMI_FN_MOD = 2, BCMI_FN_AND = 3,
    BCMI_FN_SIGNBIT = 4
};
enum bcmi_cmp {
    BCMI_CMP_LT = 0, BCMI_CMP_LE = 1, BCMI_CMP_GT = 2, BCMI_CMP_GE = 3,
    BCMI_CMP_EQ = 4, BCMI_CMP_NE = 5
};
typedef struct bcmi_predicate {
    int32_t fn;       /* bcmi_pred_fn */
    int32_t cmp;      /* bcmi_cmp */
    int32_t negate;   /* != 0: not1(predicate) */
    int32_t reserved; /* 0 */
    uint64_t operand; /* a */
    uint64_t value;   /* c */
} bcmi_predicate;

/* what bcmi_copy_if writes for each selected element */
enum bcmi_select_out {
    BCMI_SELECT_VALUES = 0,    /* copy_if (copy_if.hpp:46-62) */
    BCMI_SELECT_INDEX_U32 = 1, /* detail::copy_index_if (copy_if.hpp:25-41) */
    BCMI_SELECT_INDEX_U64 = 2
};

/* copy_if / remove_if / remove (copy_if.hpp:46-62 via transform_if.hpp:
 * 33-82; remove_if.hpp, remove.hpp are copy_if of not1(pred) back into the
 * range): the elements (or their indices) for which pred holds, in input
 * order.  out == in is allowed (the in-place remove_if).  *count_out
 * (device pointer, uint64) receives the number written; the shim reads it
 * back (blocking, as transform_if.hpp:57-62 does).  n < 2^32. */
int bcmi_copy_if(bcmi_queue queue, int dtype, const void *in, void *out,
                 size_t n, const bcmi_predicate *pred, int out_kind,
                 uint64_t *count_out);

/* count_if / count (count_if.hpp, count.hpp: count(v) is count_if(_1 == v)):
 * *count_out (device, uint64) = number of elements for which pred holds. */
int bcmi_count_if(bcmi_queue queue, int dtype, const void *in, size_t n,
                  const bcmi_predicate *pred, uint64_t *count_out);

/* unique_copy / unique with equal_to (unique_copy.hpp:99-175, unique.hpp):
 * the first element of each run of adjacent equal elements (float ==, so
 * -0 and +0 are one run and every NaN is kept).  out == in is allowed (the
 * in-place unique); elements past the count are left as they were.
 * *count_out (device, uint64) = number written.  n < 2^32. */
int bcmi_unique(bcmi_queue queue, int dtype, const void *in, void *out,
                size_t n, uint64_t *count_out);

/* partition_copy (partition_copy.hpp): out_false != NULL -- the elements
 * for which pred holds to out_true, the others to out_false, both in input
 * order.  stable_partition / partition (stable_partition.hpp,
 * partition.hpp): out_false == NULL -- both groups into out_true, the
 * selected first (out_true == in is allowed).  *count_out (device, uint64)
 * = number of elements for which pred holds.  n < 2^32. */
int bcmi_partition(bcmi_queue queue, int dtype, const void *in,
                   void *out_true, void *out_false, size_t n,
                   const bcmi_predicate *pred, uint64_t *count_out);

/* ------------------------------- transform_reduce / inner_product */
/* An input sequence: a device array, or a counting sequence
 * (make_counting_iterator, iterator/counting_iterator.hpp: element i is
 * start + i, computed in the element type). */
typedef struct bcmi_seq {
    const void *ptr; /* device pointer of element 0; NULL: a counting sequence */
    int64_t start;   /* first value of a counting sequence */
} bcmi_seq;

/* The element maps the dispatch recognises.  The reference feeds reduce
 * through transform_iterator / zip_iterator with any function object
 * (transform_reduce.hpp, inner_product.hpp); here
 *     binary (b given):  y = g(a[i] op b[i]),  op = BCMI_PLUS, BCMI_MINUS or
 *                        BCMI_MULTIPLIES  (inner_product: op = multiplies)
 *     unary  (b NULL):   y = g(a[i] - c)      (c = 0: y = g(a[i]))
 * with g = identity, abs or square, evaluated in the element type (integers
 * wrap).  c holds one element in its low bytes. */
enum bcmi_map_fn { BCMI_MAP_ID = 0, BCMI_MAP_ABS = 1, BCMI_MAP_SQUARE = 2 };
typedef struct bcmi_map {
    int32_t op;      /* binary combine (ignored when unary) */
    int32_t fn;      /* g: bcmi_map_fn */
    uint64_t offset; /* c */
} bcmi_map;

/* transform_reduce (transform_reduce.hpp) and inner_product
 * (inner_product.hpp): *out (device, one element) = init op_r y[0] op_r
 * ... op_r y[n-1] (init NULL: no init, i.e. transform_reduce), with op_r =
 * reduce_op (plus, multiplies, min, max).  Float plus with
 * BCMI_PRECISION_COMPENSATED sums in error-free pairs and keeps the
 * rounding error of a final product or square exactly (fma), with an FP64
 * combine across blocks; out_f64 (may be NULL) receives that FP64 value.
 * The reference accumulates a float inner product in one work-item
 * (accumulate.hpp:62-69); this is parallel (divergence A5). */
int bcmi_transform_reduce(bcmi_queue queue, int dtype, const bcmi_seq *a,
                          const bcmi_seq *b, size_t n, const bcmi_map *map,
                          int reduce_op, const void *init, void *out,
                          double *out_f64, int precision);

/* ------------------------------------------------- fancy-iterator inputs */
/* An input range as the reference's iterators present it:
 *   BCMI_ITER_BUFFER    element i = ptr[i]           (buffer_iterator)
 *   BCMI_ITER_COUNTING  element i = value + i        (counting_iterator,
 *                                                    iterator/counting_iterator.hpp)
 *   BCMI_ITER_CONSTANT  element i = value            (constant_iterator,
 *                                                    iterator/constant_iterator.hpp)
 * then y = g(element - offset) with g = fn: identity, abs or square
 * (transform_iterator over _1 - c, _1 + c (offset -c), abs<T>, _1 * _1,
 * pow(_1, 2), pown(_1, 2), pow(_1 - c, 2); iterator/transform_iterator.hpp),
 * or, for fn = BCMI_MAP_CMP + a bcmi_cmp code, y = (element cmp offset) ?
 * 1 : 0 (_1 > c and the other comparisons, which OpenCL C evaluates to an
 * int).  value and offset hold one element of the range's type in their
 * low bytes; arithmetic is in that type (integers wrap). */
enum bcmi_iter_kind { BCMI_ITER_BUFFER = 0, BCMI_ITER_COUNTING = 1, BCMI_ITER_CONSTANT = 2 };
typedef struct bcmi_iter {
    const void *ptr;
    int32_t kind;
    int32_t fn;
    uint64_t value;
    uint64_t offset;
} bcmi_iter;

/* exclusive_scan / inclusive_scan over an iterator range
 * (test/test_scan.cpp:140-286).  width 1: dtype is I32, U32, I64, U64, F32
 * or F64 with any kind and map (a plain array goes to bcmi_scan).  width 2
 * or 4: OpenCL vector elements (int2_, uint4_, float4_ ...) of dtype I32,
 * U32 or F32 from an array, op BCMI_PLUS, added component by component;
 * init then points at one vector element. */
int bcmi_scan_iter(bcmi_queue queue, int dtype, int width, int op,
                   const bcmi_iter *in, void *out, size_t n,
                   const void *init, int exclusive, int precision);

/* transform over iterator ranges (transform.hpp:38-82 reading
 * transform_iterator / counting_iterator / constant_iterator /
 * zip_iterator inputs): out[i] = a_i op b_i (op: plus, minus, multiplies,
 * divides, min, max), or out[i] = a_i when b is NULL (copy of a generated
 * or mapped range).  width 2 or 4 with dtype F32, b NULL and a->fn ==
 * BCMI_MAP_LENGTH: out[i] = length(a[i]), the Euclidean length of vector
 * element i (length<float4_>, functional/geometry.hpp). */
#define BCMI_MAP_LENGTH 3
#define BCMI_MAP_CMP 8
int bcmi_transform_iter(bcmi_queue queue, int dtype, int width, int op,
                        const bcmi_iter *a, const bcmi_iter *b, void *out,
                        size_t n);

/* dst[i] = (dst type) src[i]: the element conversion of copy() between
 * vectors of different element types (detail/copy_on_device.hpp:59-92) and
 * of a scan whose output type differs from its input type
 * (detail/scan_on_gpu.hpp:196-205: the scan computes in the input type and
 * converts on store).  OpenCL C conversions: integers wrap, int -> float
 * rounds to nearest, float -> int truncates toward zero. */
int bcmi_convert(bcmi_queue queue, int src_dtype, int dst_dtype,
                 const void *src, void *dst, size_t n);

/* The key extraction of experimental::sort_by_transform
 * (experimental/sort_by_transform.hpp): keys[i] = f(field of element i),
 * where element i is record_bytes bytes at records + i * record_bytes, the
 * field is one key_dtype value at byte key_offset (get<N> of an OpenCL
 * vector type: offset N * sizeof(component); 0 for a scalar), and f is
 * BCMI_MAP_ID or BCMI_MAP_ABS (abs<T>; integer abs wraps, so INT_MIN stays
 * INT_MIN).  The reference then sorts the elements by these keys with
 * sort_by_key; bcmi_sort(api = BCMI_SORT_BY_KEY) with the records as
 * values does that. */
int bcmi_extract_keys(bcmi_queue queue, int key_dtype, int fn,
                      const void *records, size_t record_bytes,
                      size_t key_offset, size_t n, void *keys);

/* ------------------------------------------------ comparator merge sort */
/* Comparison sort for comparators other than less<T> / greater<T> on a
 * fundamental key: the reference's merge_sort_on_gpu
 * (algorithm/detail/merge_sort_on_gpu.hpp:467-571), reached from sort
 * (sort.hpp:84-104), sort_by_key (sort_by_key.hpp:88-106), stable_sort
 * (stable_sort.hpp:36-50) and stable_sort_by_key with a user comparator.
 *
 * The comparator is  key(a, lhs_offset) CMP key(b, rhs_offset),  CMP being
 * < (greater == 0) or > (greater != 0), over fields of scalar type `dtype`
 * at byte offsets inside each record_bytes-byte record:
 *   BCMI_KEY_FIELD     the field                          a.x < b.x, a.s0 < b.s3
 *   BCMI_KEY_ABS       abs(field): integer abs is the     abs(a) < abs(b)
 *                      unsigned magnitude (abs(INT_MIN) = 2^31, as OpenCL's
 *                      abs returns an unsigned type); fabs for floats
 *   BCMI_KEY_ABS_SUM2  abs(field + field at offset2), the sum in C's
 *                      promoted type (narrow integers add in int, 32- and
 *                      64-bit integers wrap)                 abs(a.x + a.y)
 *   BCMI_KEY_LENGTH    Euclidean length of `count` consecutive float or
 *                      double fields, compared through the sum of squares
 *                      in double                             length(a) < length(b)
 * ABS_SUM2 and LENGTH need lhs_offset == rhs_offset.  Floats compare as C
 * does (-0 == +0); NaN fields give an unspecified order, as a comparator
 * that is not a strict weak order does in the reference.
 *
 * The sort is stable (ties keep their input order), which is also a valid
 * result of the reference's unstable merge sort.  values (may be NULL) are
 * value_bytes-byte records moved with their records.  Records are 1, 2, 4,
 * 8, 16, 32 or 64 bytes; n < 2^32.  Temporaries come from the queue's
 * scratch arena. */
typedef enum {
    BCMI_KEY_FIELD = 0,
    BCMI_KEY_ABS = 1,
    BCMI_KEY_ABS_SUM2 = 2,
    BCMI_KEY_LENGTH = 3
} bcmi_key_form;

typedef struct bcmi_comparator {
    int32_t dtype;        /* bcmi_dtype of the compared fields */
    int32_t form;         /* bcmi_key_form */
    uint32_t lhs_offset;  /* field offset read from the left operand */
    uint32_t rhs_offset;  /* field offset read from the right operand */
    uint32_t offset2;     /* ABS_SUM2: offset of the second addend */
    uint32_t count;       /* LENGTH: number of components (1..16) */
    int32_t greater;      /* 0: key(a) < key(b); otherwise key(a) > key(b) */
} bcmi_comparator;

int bcmi_merge_sort(bcmi_queue queue, const bcmi_comparator *cmp,
                    void *records, size_t record_bytes, void *values,
                    size_t value_bytes, size_t n);

/* --------------------------------------------------------- hipGraphs */
/* Capture everything enqueued on the queue between begin and end into a
 * replayable graph (queue scratch must be reserved beforehand). */
typedef struct bcmi_graph_s *bcmi_graph;
int bcmi_graph_begin(bcmi_queue queue);
int bcmi_graph_end(bcmi_queue queue, bcmi_graph *graph);
int bcmi_graph_launch(bcmi_graph graph, bcmi_queue queue);
int bcmi_graph_destroy(bcmi_graph graph);

/* ------------------------------------------------------------ tuning */
/* parameter_cache (detail/parameter_cache.hpp:63-82): per-device
 * (key, param) -> value, persisted as JSON.  The first bcmi_queue_create
 * loads the file named by the environment variable BCMI_TUNING_CACHE, else
 * tuning/gfx950.json beside libbcmi.so when present, without overriding
 * values already set with bcmi_tune_set. */
int bcmi_tune_get(const char *key, const char *param, unsigned default_value,
                  unsigned *value);
int bcmi_tune_set(const char *key, const char *param, unsigned value);
int bcmi_tuning_cache_load(const char *path);
int bcmi_tuning_cache_save(const char *path);

/* ------------------------------------------------------------ profiling */
/* When enabled, the queue brackets each kernel launch with HIP events on
 * its own stream; bcmi_profile_query sums them per kernel name (blocking). */
int bcmi_profile_enable(bcmi_queue queue, int enable);
int bcmi_profile_query(bcmi_queue queue, const char *kernel, int *launches,
                       double *total_ms);
int bcmi_profile_reset(bcmi_queue queue);

#ifdef __cplusplus
}
#endif
#endif /* BCMI_H */
