/* cpu_column.c -- multi-threaded restatement of the reference's CPU-device
 * column (what Boost.Compute runs on a CPU OpenCL device such as PoCL),
 * used ONLY as the timed CPU baseline of the benchmark.  Test
 * infrastructure: nothing in the library links or calls it.
 *
 * PoCL runs each work-item range over the host cores; here the same work
 * decomposition runs on OpenMP threads (`threads` = the compute units the
 * reference would see, CL_DEVICE_MAX_COMPUTE_UNITS):
 *   sort          merge_sort_on_cpu (merge_sort_on_cpu.hpp:236-299): stable
 *                 insertion sort of 64-element blocks (:161-222), then
 *                 log2(n/64) merge passes, one work-item per pair of blocks
 *                 (merge_blocks :26-112); once at most 8 blocks are left and
 *                 n >= 2^21, each pair is merged by merge path in tiles of
 *                 1024 outputs (dispatch_merge_blocks :128-158,
 *                 merge_with_merge_path.hpp:140-181)
 *   sort_by_key   merge_sort_by_key_on_cpu (:301-360): the same blocks and
 *                 pair merges, never merge path
 *   scan          scan_on_cpu (scan_on_cpu.hpp:28-199): cu+1 chunks, pass 1
 *                 folds chunk totals, pass 2 scans each chunk from the fold
 *                 of the earlier totals
 *   reduce        reduce_on_cpu (reduce_on_cpu.hpp:29-104): cu chunks folded
 *                 in parallel, then the partials serially
 *   accumulate    float plus: one serial left fold (accumulate.hpp:102-122,
 *                 serial_accumulate.hpp:22-50)
 *   reduce_by_key serial_reduce_by_key (serial_reduce_by_key.hpp:28-100): one
 *                 work-item
 *   transform     copy_on_device_cpu (copy_on_device.hpp:63-92): cu chunks
 * Comparisons are operator< on the native type (-0.0 == +0.0, stable). */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BLOCK 64

/* ------------------------------------------------------------ merge sort */
#define DEFINE_MERGE_SORT(NAME, K, V, HAS_V)                                                         \
static void NAME##_insertion(K *k, V *v, size_t lo, size_t hi)                                   \
{                                                                                                \
    for (size_t i = lo + 1; i < hi; ++i) {                                                       \
        K key = k[i];                                                                            \
        V val;                                                                                   \
        if (HAS_V) val = v[i];                                                                   \
        size_t pos = i;                                                                          \
        while (pos > lo && key < k[pos - 1]) {                                                   \
            k[pos] = k[pos - 1];                                                                 \
            if (HAS_V) v[pos] = v[pos - 1];                                                      \
            pos--;                                                                               \
        }                                                                                        \
        k[pos] = key;                                                                            \
        if (HAS_V) v[pos] = val;                                                                 \
    }                                                                                            \
}                                                                                                \
/* stable merge of [a0,a1) and [b0,b1) of ks/vs into kd/vd from o */                           \
static void NAME##_merge(const K *ks, const V *vs, K *kd, V *vd, size_t a0, size_t a1, size_t b0, \
                         size_t b1, size_t o)                                                    \
{                                                                                                \
    while (a0 < a1 && b0 < b1) {                                                                 \
        if (ks[b0] < ks[a0]) {                                                                   \
            kd[o] = ks[b0];                                                                      \
            if (HAS_V) vd[o] = vs[b0];                                                           \
            b0++;                                                                                \
        } else {                                                                                 \
            kd[o] = ks[a0];                                                                      \
            if (HAS_V) vd[o] = vs[a0];                                                           \
            a0++;                                                                                \
        }                                                                                        \
        o++;                                                                                     \
    }                                                                                            \
    for (; a0 < a1; ++a0, ++o) {                                                                 \
        kd[o] = ks[a0];                                                                          \
        if (HAS_V) vd[o] = vs[a0];                                                               \
    }                                                                                            \
    for (; b0 < b1; ++b0, ++o) {                                                                 \
        kd[o] = ks[b0];                                                                          \
        if (HAS_V) vd[o] = vs[b0];                                                               \
    }                                                                                            \
}                                                                                                \
/* merge path: split point on diagonal d of A=[a0,a1), B=[b0,b1): i from A */                  \
static size_t NAME##_diag(const K *ks, size_t a0, size_t na, size_t b0, size_t nb, size_t d)      \
{                                                                                                \
    size_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;                                       \
    while (lo < hi) {                                                                            \
        size_t i = (lo + hi) / 2;                                                                \
        /* take A[i] before B[d-1-i] unless B[d-1-i] < A[i] */                                   \
        if (ks[b0 + d - 1 - i] < ks[a0 + i]) hi = i;                                             \
        else lo = i + 1;                                                                         \
    }                                                                                            \
    return lo;                                                                                   \
}                                                                                                \
void NAME(K *keys, V *vals, size_t n, int threads, int merge_path)                               \
{                                                                                                \
    if (n < 2) return;                                                                           \
    omp_set_num_threads(threads);                                                                \
    if (n <= 512) {                                                                              \
        NAME##_insertion(keys, vals, 0, n);                                                      \
        return;                                                                                  \
    }                                                                                            \
    const long nblocks = (long)((n + BLOCK - 1) / BLOCK);                                        \
    _Pragma("omp parallel for schedule(static)")                                                \
    for (long b = 0; b < nblocks; ++b) {                                                         \
        size_t lo = (size_t)b * BLOCK, hi = lo + BLOCK < n ? lo + BLOCK : n;                     \
        NAME##_insertion(keys, vals, lo, hi);                                                    \
    }                                                                                            \
    K *kt = (K *)malloc(n * sizeof(K));                                                          \
    V *vt = HAS_V ? (V *)malloc(n * sizeof(V)) : NULL;                                           \
    K *ks = keys, *kd = kt;                                                                      \
    V *vs = vals, *vd = vt;                                                                      \
    for (size_t w = BLOCK; w < n; w *= 2) {                                                      \
        const size_t blocks = (n + w - 1) / w;                                                   \
        if (merge_path && blocks <= 8 && n >= (1u << 21)) {                                      \
            for (size_t lo = 0; lo < n; lo += 2 * w) {                                           \
                const size_t a1 = lo + w < n ? lo + w : n, b1 = lo + 2 * w < n ? lo + 2 * w : n; \
                const size_t na = a1 - lo, nb = b1 - a1, tot = na + nb;                          \
                const long tiles = (long)((tot + 1023) / 1024);                                  \
                _Pragma("omp parallel for schedule(static)")                                    \
                for (long t = 0; t < tiles; ++t) {                                               \
                    size_t d0 = (size_t)t * 1024, d1 = d0 + 1024 < tot ? d0 + 1024 : tot;        \
                    size_t i0 = NAME##_diag(ks, lo, na, a1, nb, d0);                             \
                    size_t i1 = NAME##_diag(ks, lo, na, a1, nb, d1);                             \
                    NAME##_merge(ks, vs, kd, vd, lo + i0, lo + i1, a1 + d0 - i0, a1 + d1 - i1, lo + d0); \
                }                                                                                \
            }                                                                                    \
        } else {                                                                                 \
            const long pairs = (long)((n + 2 * w - 1) / (2 * w));                                \
            _Pragma("omp parallel for schedule(dynamic, 1)")                                    \
            for (long p = 0; p < pairs; ++p) {                                                   \
                size_t lo = (size_t)p * 2 * w;                                                   \
                size_t a1 = lo + w < n ? lo + w : n, b1 = lo + 2 * w < n ? lo + 2 * w : n;       \
                NAME##_merge(ks, vs, kd, vd, lo, a1, a1, b1, lo);                                \
            }                                                                                    \
        }                                                                                        \
        K *tk = ks; ks = kd; kd = tk;                                                            \
        V *tv = vs; vs = vd; vd = tv;                                                            \
    }                                                                                            \
    if (ks != keys) {                                                                            \
        memcpy(keys, ks, n * sizeof(K));                                                         \
        if (HAS_V) memcpy(vals, vs, n * sizeof(V));                                              \
    }                                                                                            \
    free(kt);                                                                                    \
    free(vt);                                                                                    \
}

DEFINE_MERGE_SORT(cc_sort_u32, uint32_t, uint32_t, 0)
DEFINE_MERGE_SORT(cc_sort_by_key_f32_u32, float, uint32_t, 1)
DEFINE_MERGE_SORT(cc_sort_by_key_u64_u32, uint64_t, uint32_t, 1)

/* one stable_sort_by_key call per segment (the per-call baseline of the
 * segmented sort); each call runs its own passes on all threads */
void cc_segmented_sort_u64_u32(uint64_t *keys, uint32_t *vals, const uint64_t *offsets, size_t nseg, int threads)
{
    for (size_t s = 0; s < nseg; ++s)
        cc_sort_by_key_u64_u32(keys + offsets[s], vals + offsets[s], (size_t)(offsets[s + 1] - offsets[s]),
                               threads, 0);
}

/* ------------------------------------------------------------------ scan */
/* int32 plus (wrapping), cu = threads work-items over cu+1 chunks */
void cc_scan_i32(const int32_t *x, int32_t *out, size_t n, int exclusive, int32_t init, int threads)
{
    omp_set_num_threads(threads);
    const size_t G = (size_t)threads;
    if (n < 16384 * sizeof(int32_t) || G < 2) {
        uint32_t acc = (uint32_t)init;
        for (size_t i = 0; i < n; ++i) {
            if (exclusive) {
                out[i] = (int32_t)acc;
                acc += (uint32_t)x[i];
            } else {
                acc = i == 0 ? (uint32_t)x[0] : acc + (uint32_t)x[i];
                out[i] = (int32_t)acc;
            }
        }
        return;
    }
    const size_t block = (n + G) / (G + 1);
    uint32_t *part = (uint32_t *)calloc(G + 1, sizeof(uint32_t));
    /* pass 1: totals of chunks 0..G-1; chunk 0 is written here */
    _Pragma("omp parallel for schedule(static)")
    for (long g = 0; g < (long)G; ++g) {
        size_t lo = (size_t)g * block, hi = lo + block < n ? lo + block : n;
        uint32_t acc = (g == 0 && exclusive) ? (uint32_t)init : 0u;
        for (size_t i = lo; i < hi; ++i) {
            if (g == 0) {
                if (exclusive) {
                    out[i] = (int32_t)acc;
                    acc += (uint32_t)x[i];
                } else {
                    acc += (uint32_t)x[i];
                    out[i] = (int32_t)acc;
                }
            } else {
                acc += (uint32_t)x[i];
            }
        }
        part[g] = acc;
    }
    /* pass 2: chunk g+1 from the fold of totals 0..g */
    _Pragma("omp parallel for schedule(static)")
    for (long g = 0; g < (long)G; ++g) {
        uint32_t seed = 0;
        for (long j = 0; j <= g; ++j) seed += part[j];
        size_t lo = (size_t)(g + 1) * block, hi = lo + block < n ? lo + block : n;
        for (size_t i = lo; i < hi; ++i) {
            if (exclusive) {
                out[i] = (int32_t)seed;
                seed += (uint32_t)x[i];
            } else {
                seed += (uint32_t)x[i];
                out[i] = (int32_t)seed;
            }
        }
    }
    free(part);
}

/* ---------------------------------------------------------------- reduce */
float cc_reduce_f32(const float *x, size_t n, int threads)
{
    omp_set_num_threads(threads);
    const size_t cu = (size_t)threads;
    if (n < 16384 * sizeof(float) || n < cu) {
        float r = x[0];
        for (size_t i = 1; i < n; ++i) r = r + x[i];
        return r;
    }
    float *part = (float *)malloc(cu * sizeof(float));
    const size_t block = (size_t)ceilf((float)n / (float)cu);
    _Pragma("omp parallel for schedule(static)")
    for (long g = 0; g < (long)cu; ++g) {
        size_t lo = (size_t)g * block, hi = lo + block < n ? lo + block : n;
        float r = lo < n ? x[lo] : 0.0f;
        for (size_t i = lo + 1; i < hi; ++i) r = r + x[i];
        part[g] = r;
    }
    float r = part[0];
    for (size_t g = 1; g < cu; ++g) r = r + part[g];
    free(part);
    return r;
}

float cc_accumulate_f32(const float *x, size_t n, float init)
{
    float r = init;
    for (size_t i = 0; i < n; ++i) r = r + x[i];
    return r;
}

/* --------------------------------------------------------- reduce_by_key */
size_t cc_reduce_by_key_i32_f32(const int32_t *keys, const float *vals, size_t n, int32_t *ko, float *vo)
{
    if (!n) return 0;
    size_t r = 0;
    float acc = vals[0];
    for (size_t i = 1; i < n; ++i) {
        if (keys[i] == keys[i - 1]) {
            acc = acc + vals[i];
        } else {
            ko[r] = keys[i - 1];
            vo[r++] = acc;
            acc = vals[i];
        }
    }
    ko[r] = keys[n - 1];
    vo[r++] = acc;
    return r;
}

/* ----------------------------------------------------------------- saxpy */
void cc_saxpy_f32(float a, const float *x, const float *y, float *out, size_t n, int threads)
{
    omp_set_num_threads(threads);
    const size_t cu = (size_t)threads, block = (n + cu - 1) / cu;
    _Pragma("omp parallel for schedule(static)")
    for (long g = 0; g < (long)cu; ++g) {
        size_t lo = (size_t)g * block, hi = lo + block < n ? lo + block : n;
        for (size_t i = lo; i < hi; ++i) out[i] = fmaf(a, x[i], y[i]);
    }
}
