/*
 * bcmi_oracle.h -- CPU oracle for the data-parallel primitive core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  Nothing in the product library (compute_amd/) links or calls it.
 *
 * Every function restates the reference's algorithm (Boost.Compute 0.5.0,
 * include/boost/compute/...) in plain C; each definition cites the file:line
 * it follows.  Float helpers that have no reference counterpart (double-double
 * sums, used as the FP64 oracle) say so.
 *
 * How each function is pinned (known-answer tests, the reference's own golden
 * vectors, closed forms, brute force) is listed in tests/test_oracle.py.
 */
#ifndef BCMI_ORACLE_H
#define BCMI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- RNG -- */
typedef struct { uint32_t mt[624]; int idx; } orc_mt32;
typedef struct { uint64_t mt[312]; int idx; } orc_mt64;

void     orc_mt32_seed(orc_mt32 *g, uint32_t seed);
uint32_t orc_mt32_next(orc_mt32 *g);
void     orc_mt64_seed(orc_mt64 *g, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *g);
/* n raw outputs of std::mt19937(seed) / std::mt19937_64(seed) */
void orc_mt32_fill(uint32_t seed, uint32_t *out, size_t n);
void orc_mt64_fill(uint64_t seed, uint64_t *out, size_t n);

/* ------------------------------------------------ workload generators -- */
void   orc_gen_sort_u32(uint32_t seed, uint32_t *keys, size_t n);
void   orc_gen_sort_by_key_f32(float *keys, uint32_t *vals, size_t n);
void   orc_gen_scan_i32(int32_t *x, size_t n);
void   orc_gen_saxpy_f32(float *x, float *y, size_t n);
void   orc_gen_lognormal_f32(float *x, size_t n);
size_t orc_gen_segments(uint64_t *offsets, size_t nseg, size_t max_len);
void   orc_gen_segment_data(uint64_t *keys, uint32_t *vals, size_t n);
size_t orc_gen_runs_i32_f32(int32_t *keys, float *vals, size_t n, size_t max_run);

/* --------------------------------------------------------- sort paths -- */
/* kind: 0 unsigned, 1 signed, 2 floating point (IEEE binary32/64) */
uint64_t orc_radix_key(uint64_t bits, int kbytes, int kind, int descending);
/* stable sort by the radix key: the GPU column of sort/sort_by_key/stable_sort */
void orc_radix_order_sort(void *keys, void *vals, size_t n, int kbytes,
                          int kind, int descending, int vbytes);
/* one-work-item insertion sort with operator< / operator> on the native type */
void orc_insertion_sort(void *keys, void *vals, size_t n, int kbytes,
                        int kind, int descending, int vbytes);
/* public sort/sort_by_key/stable_sort/stable_sort_by_key dispatch (GPU column)
 * api: 0 sort, 1 sort_by_key, 2 stable_sort, 3 stable_sort_by_key */
void orc_sort_dispatch(int api, void *keys, void *vals, size_t n, int kbytes,
                       int kind, int descending, int vbytes);
/* stable comparison sort (merge sort, operator<): CPU-device column */
void orc_compare_stable_sort(void *keys, void *vals, size_t n, int kbytes,
                             int kind, int descending, int vbytes);
void orc_segmented_sort(void *keys, void *vals, const uint64_t *offsets,
                        size_t nseg, int kbytes, int kind, int descending,
                        int vbytes);

/* --------------------------------------------------------------- scans -- */
/* dtype codes shared with the product C-ABI */
enum { ORC_I8 = 1, ORC_U8, ORC_I16, ORC_U16, ORC_I32, ORC_U32, ORC_I64, ORC_U64,
       ORC_F32, ORC_F64 };
enum { ORC_PLUS = 0, ORC_MULTIPLIES = 1, ORC_MIN = 2, ORC_MAX = 3 };
/* sequential left-fold scan in the element type (wrapping for integers) */
void orc_scan(int dtype, int op, const void *in, void *out, size_t n,
              const void *init, int exclusive);
/* float scans: double-double accurate prefix sums rounded to double */
void orc_scan_f32_dd(const float *in, double *out, size_t n, int exclusive,
                     double init);
/* segmented (by adjacent-equal int32 keys) exclusive/inclusive plus-scan */
void orc_segmented_scan_i32keys(const int32_t *keys, int vdtype,
                                const void *vals, void *out, size_t n,
                                int exclusive, const void *init);
void orc_segmented_scan_f32_dd(const int32_t *keys, const float *vals,
                               double *out, size_t n, int exclusive);

/* ------------------------------------------------------------- reduce -- */
void   orc_reduce(int dtype, int op, const void *in, size_t n, void *out);
double orc_sum_f32_dd(const float *x, size_t n, double *lo);
double orc_sum_f64_dd(const double *x, size_t n, double *lo);
float  orc_reduce_gpu_order_f32(const float *x, size_t n);
float  orc_reduce_cpu_order_f32(const float *x, size_t n, unsigned cu);
float  orc_accumulate_f32(const float *x, size_t n, float init);
/* scan_on_cpu (CPU-device column) association with cu compute units, plus
 * on ORC_F32 or ORC_I32; out must be pre-filled by the caller where the
 * reference leaves it unwritten (cu < 2).  0 ok, -1 undefined (empty chunk),
 * -2 bad dtype. */
int    orc_scan_cpu_order(int dtype, const void *in, void *out, size_t n,
                          int exclusive, const void *init, unsigned cu);
float  orc_scan_gpu_order_f32_last(const float *x, size_t n);
void   orc_scan_gpu_order_f32(const float *x, float *out, size_t n, int exclusive);

/* ------------------------------------------------------ reduce_by_key -- */
size_t orc_reduce_by_key(int kdtype, int vdtype, int op, const void *keys,
                         const void *vals, size_t n, void *keys_out,
                         void *vals_out);
size_t orc_reduce_by_key_f32_dd(const int32_t *keys, const float *vals,
                                size_t n, int32_t *keys_out, double *sums);

/* ---------------------------------------------------------- transform -- */
void orc_saxpy_f32(float a, const float *x, const float *y, float *out,
                   size_t n);
void orc_transform(int dtype, int op, const void *a, const void *b, void *out,
                   size_t n);

/* ------------------------------------------- selection (stream compaction) --
 * A unary predicate "f(_1) cmp c" (f = _1, _1*a, _1%a, _1&a, signbit(_1);
 * cmp < <= > >= == !=; negate = not1), evaluated with C's usual arithmetic
 * conversions: 8/16-bit integers promote to int, signbit yields int. */
enum { ORC_FN_ID = 0, ORC_FN_MUL = 1, ORC_FN_MOD = 2, ORC_FN_AND = 3, ORC_FN_SIGNBIT = 4 };
enum { ORC_CMP_LT = 0, ORC_CMP_LE = 1, ORC_CMP_GT = 2, ORC_CMP_GE = 3, ORC_CMP_EQ = 4, ORC_CMP_NE = 5 };
typedef struct {
    int32_t fn, cmp, negate, reserved;
    uint64_t operand, value;   /* low bytes: one value of the comparison type */
} orc_predicate;

int    orc_pred_eval(int dtype, const void *in, size_t i, const orc_predicate *p);
/* out_kind 0: values, 1: uint32 indices, 2: uint64 indices */
size_t orc_copy_if(int dtype, const void *in, size_t n, const orc_predicate *p,
                   int out_kind, void *out);
size_t orc_count_if(int dtype, const void *in, size_t n, const orc_predicate *p);
size_t orc_unique_copy(int dtype, const void *in, size_t n, void *out);
size_t orc_partition_copy(int dtype, const void *in, size_t n,
                          const orc_predicate *p, void *out_true, void *out_false);
size_t orc_stable_partition(int dtype, void *data, size_t n, const orc_predicate *p);

/* --------------------------------------- transform_reduce / inner_product --
 * y_i = g(a_i op b_i) (binary) or g(a_i - c) (unary), g = id | abs | square,
 * op = plus | minus | multiplies (the ORC_ op codes, minus = 4), each step in
 * the element type.  A NULL array is a counting sequence start + i. */
enum { ORC_MAP_ID = 0, ORC_MAP_ABS = 1, ORC_MAP_SQUARE = 2 };
typedef struct {
    int32_t op, fn;
    uint64_t offset;
} orc_map;

/* init op_r y_0 op_r ... (left fold in the element type; init NULL: none) */
void   orc_map_reduce(int dtype, const void *a, int64_t a_start, const void *b, int64_t b_start, int binary,
                      size_t n, const orc_map *m, int rop, const void *init, void *out);
/* floats: sum of the terms with each final product or square exact, in
 * double-double; returns hi, *lo the low part */
double orc_map_sum_dd(int dtype, const void *a, int64_t a_start, const void *b, int64_t b_start, int binary,
                      size_t n, const orc_map *m, double *lo);

#ifdef __cplusplus
}
#endif
#endif
