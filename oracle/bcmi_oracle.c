/*
 * bcmi_oracle.c -- CPU oracle (test infrastructure only; see bcmi_oracle.h).
 *
 * Reference = Boost.Compute 0.5.0.  Paths below are relative to its root.
 * Built with -ffp-contract=off so every float expression rounds exactly as
 * written (the reference's kernels are restated operation by operation).
 */
#include "bcmi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Mersenne twister (the published MT19937 / MT19937-64 algorithms of
 * Matsumoto & Nishimura, as specified for std::mt19937 / std::mt19937_64
 * in C++11 [rand.eng.mers]).  Pinned by the standard's known answers:
 * the 10000th output of a default-seeded (5489) engine.
 * ==================================================================== */
void orc_mt32_seed(orc_mt32 *g, uint32_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

uint32_t orc_mt32_next(orc_mt32 *g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void orc_mt64_seed(orc_mt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ull * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}

uint64_t orc_mt64_next(orc_mt64 *g)
{
    if (g->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t y = (g->mt[i] & 0xFFFFFFFF80000000ull) |
                         (g->mt[(i + 1) % 312] & 0x7FFFFFFFull);
            g->mt[i] = g->mt[(i + 156) % 312] ^ (y >> 1) ^
                       ((y & 1ull) ? 0xB5026F5AA96619E9ull : 0ull);
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ull;
    y ^= (y << 17) & 0x71D67FFFEDA60000ull;
    y ^= (y << 37) & 0xFFF7EEE000000000ull;
    y ^= y >> 43;
    return y;
}

void orc_mt32_fill(uint32_t seed, uint32_t *out, size_t n)
{
    orc_mt32 g;
    orc_mt32_seed(&g, seed);
    for (size_t i = 0; i < n; ++i) out[i] = orc_mt32_next(&g);
}

void orc_mt64_fill(uint64_t seed, uint64_t *out, size_t n)
{
    orc_mt64 g;
    orc_mt64_seed(&g, seed);
    for (size_t i = 0; i < n; ++i) out[i] = orc_mt64_next(&g);
}

/* ======================================================================
 * Workload generators.  Explicit transforms of raw engine output (never a
 * std::*_distribution, whose algorithms are implementation-defined).
 * The product's workload generator (compute_amd/csrc/workloads.cpp) is an
 * independent implementation of the same definitions; the tests compare the
 * two.
 * ==================================================================== */
static const double TWO_M32 = 2.3283064365386962890625e-10; /* 2^-32 */
static const double TWO_M53 = 1.1102230246251565404236316680908203125e-16;
static const double TWO_PI = 6.283185307179586476925286766559;

/* Box-Muller, cosine branch, two 32-bit draws: u1 in (0,1], u2 in [0,1) */
static double bm32(orc_mt32 *g)
{
    double u1 = ((double)orc_mt32_next(g) + 1.0) * TWO_M32;
    double u2 = (double)orc_mt32_next(g) * TWO_M32;
    return sqrt(-2.0 * log(u1)) * cos(TWO_PI * u2);
}

static double bm64(orc_mt64 *g)
{
    double u1 = ((double)(orc_mt64_next(g) >> 11) + 1.0) * TWO_M53;
    double u2 = (double)(orc_mt64_next(g) >> 11) * TWO_M53;
    return sqrt(-2.0 * log(u1)) * cos(TWO_PI * u2);
}

/* uint32 keys uniform over the full range: key[i] = gen() */
void orc_gen_sort_u32(uint32_t seed, uint32_t *keys, size_t n)
{
    orc_mt32_fill(seed, keys, n);
}

/* float32 keys uniform in [-1e6, 1e6] with 1% signed zeros and 1% duplicates,
 * payload = original index. */
void orc_gen_sort_by_key_f32(float *keys, uint32_t *vals, size_t n)
{
    orc_mt32 g, g2;
    orc_mt32_seed(&g, 0xB0057002u);
    orc_mt32_seed(&g2, 0xB0057012u);
    for (size_t i = 0; i < n; ++i) {
        float u = (float)(orc_mt32_next(&g) >> 8) * 0x1p-24f;
        float t = 2e6f * u;
        float k = -1e6f + t;
        uint32_t r = orc_mt32_next(&g2);
        keys[i] = k;
        if (r % 200u == 0u) {
            keys[i] = -0.0f;
        } else if (r % 200u == 1u) {
            keys[i] = 0.0f;
        } else if (r % 100u == 2u) {
            uint32_t j = orc_mt32_next(&g2) % (uint32_t)(i + 1);
            keys[i] = keys[j];
        }
        vals[i] = (uint32_t)i;
    }
}

/* int32 uniform in [-100, 100] */
void orc_gen_scan_i32(int32_t *x, size_t n)
{
    orc_mt32 g;
    orc_mt32_seed(&g, 0xB0057003u);
    for (size_t i = 0; i < n; ++i) x[i] = (int32_t)(orc_mt32_next(&g) % 201u) - 100;
}

/* x then y, each N(0,1) (Box-Muller in double, rounded to float) */
void orc_gen_saxpy_f32(float *x, float *y, size_t n)
{
    orc_mt32 g;
    orc_mt32_seed(&g, 0xB0057004u);
    for (size_t i = 0; i < n; ++i) x[i] = (float)bm32(&g);
    for (size_t i = 0; i < n; ++i) y[i] = (float)bm32(&g);
}

/* +-exp(3 z), z ~ N(0,1); the sign is the low bit of the next draw */
void orc_gen_lognormal_f32(float *x, size_t n)
{
    orc_mt32 g;
    orc_mt32_seed(&g, 0xB0057005u);
    for (size_t i = 0; i < n; ++i) {
        double z = bm32(&g);
        double v = exp(3.0 * z);
        if (orc_mt32_next(&g) & 1u) v = -v;
        x[i] = (float)v;
    }
}

/* Segment lengths: lognormal, median 256, sigma = sqrt(2 ln 4), clamped to
 * [1, max_len]; segment 0 is forced to max_len.  offsets has nseg+1 entries.
 * Returns the total length. */
size_t orc_gen_segments(uint64_t *offsets, size_t nseg, size_t max_len)
{
    orc_mt64 g;
    orc_mt64_seed(&g, 0xB0057006ull);
    const double sigma = sqrt(2.0 * log(4.0));
    const double mu = log(256.0);
    offsets[0] = 0;
    for (size_t s = 0; s < nseg; ++s) {
        double z = bm64(&g);
        long long L = llround(exp(mu + sigma * z));
        if (L < 1) L = 1;
        if ((unsigned long long)L > max_len) L = (long long)max_len;
        if (s == 0) L = (long long)max_len;
        offsets[s + 1] = offsets[s] + (uint64_t)L;
    }
    return (size_t)offsets[nseg];
}

/* uint64 keys uniform over the full range, payload = global index */
void orc_gen_segment_data(uint64_t *keys, uint32_t *vals, size_t n)
{
    orc_mt64 g;
    orc_mt64_seed(&g, 0xB0057016ull);
    for (size_t i = 0; i < n; ++i) {
        keys[i] = orc_mt64_next(&g);
        vals[i] = (uint32_t)i;
    }
}

/* Sorted int32 keys forming runs of geometric length (p = 1/32, capped at
 * max_run); key = run index; value = float(z * 10^(6u - 3)).  Returns the
 * number of runs. */
size_t orc_gen_runs_i32_f32(int32_t *keys, float *vals, size_t n, size_t max_run)
{
    orc_mt32 g;
    orc_mt32_seed(&g, 0xB0057007u);
    const double l1p = log(1.0 - 1.0 / 32.0);
    size_t i = 0, run = 0;
    while (i < n) {
        double u = ((double)orc_mt32_next(&g) + 0.5) * TWO_M32;
        size_t L = 1 + (size_t)floor(log(u) / l1p);
        if (L > max_run) L = max_run;
        for (size_t j = 0; j < L && i < n; ++j, ++i) {
            double z = bm32(&g);
            double e = 6.0 * ((double)orc_mt32_next(&g) * TWO_M32) - 3.0;
            keys[i] = (int32_t)run;
            vals[i] = (float)(z * pow(10.0, e));
        }
        ++run;
    }
    return run;
}

/* ======================================================================
 * Sorting
 * ==================================================================== */
static uint64_t load_bits(const void *p, size_t i, int kb)
{
    switch (kb) {
    case 1: return ((const uint8_t *)p)[i];
    case 2: return ((const uint16_t *)p)[i];
    case 4: return ((const uint32_t *)p)[i];
    default: return ((const uint64_t *)p)[i];
    }
}

static void store_bits(void *p, size_t i, int kb, uint64_t v)
{
    switch (kb) {
    case 1: ((uint8_t *)p)[i] = (uint8_t)v; break;
    case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
    default: ((uint64_t *)p)[i] = v; break;
    }
}

/* radix() key transform, algorithm/detail/radix_sort.hpp:98-129.
 * Ascending: float  x ^ (-(x >> msb) | signbit)    (:102-104)
 *            signed x ^ signbit                    (:105-106)
 *            unsigned x                            (:107-108)
 * Descending: float  (-x) ^ (-(x >> msb) | signbit) (:117-119)
 *             signed (-x) ^ signbit                 (:120-121)
 *             unsigned (T)(-1) - x                  (:122-123)
 * All arithmetic is modulo 2^W on the W-bit unsigned image of the key (the
 * reference's T is radix_sort_value_type<sizeof(key)>, :49-76). */
uint64_t orc_radix_key(uint64_t bits, int kbytes, int kind, int descending)
{
    const int W = kbytes * 8;
    const uint64_t mask = (W == 64) ? ~0ull : ((1ull << W) - 1ull);
    const uint64_t sign = 1ull << (W - 1);
    const uint64_t x = bits & mask;
    uint64_t fmask = ((x >> (W - 1)) ? mask : 0ull) | sign;
    if (!descending) {
        if (kind == 2) return (x ^ fmask) & mask;
        if (kind == 1) return (x ^ sign) & mask;
        return x;
    }
    uint64_t neg = (0ull - x) & mask;
    if (kind == 2) return (neg ^ fmask) & mask;
    if (kind == 1) return (neg ^ sign) & mask;
    return (mask - x) & mask;
}

/* Stable bottom-up merge sort of an index permutation by a 64-bit key. */
static void stable_index_sort(const uint64_t *key, size_t *idx, size_t n)
{
    size_t *tmp = (size_t *)malloc(n * sizeof(size_t) + 1);
    size_t *a = idx, *b = tmp;
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n;
            size_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                /* take from the right only when strictly smaller: stable */
                if (key[a[j]] < key[a[i]]) b[o++] = a[j++];
                else b[o++] = a[i++];
            }
            while (i < mid) b[o++] = a[i++];
            while (j < hi) b[o++] = a[j++];
        }
        size_t *t = a; a = b; b = t;
    }
    if (a != idx) memcpy(idx, a, n * sizeof(size_t));
    free(tmp);
}

static void apply_perm(void *keys, void *vals, const size_t *idx, size_t n,
                       int kbytes, int vbytes)
{
    if (n == 0) return;
    void *kt = malloc(n * (size_t)kbytes);
    for (size_t i = 0; i < n; ++i)
        memcpy((char *)kt + i * kbytes, (char *)keys + idx[i] * kbytes, (size_t)kbytes);
    memcpy(keys, kt, n * (size_t)kbytes);
    free(kt);
    if (vals && vbytes > 0) {
        void *vt = malloc(n * (size_t)vbytes);
        for (size_t i = 0; i < n; ++i)
            memcpy((char *)vt + i * vbytes, (char *)vals + idx[i] * vbytes, (size_t)vbytes);
        memcpy(vals, vt, n * (size_t)vbytes);
        free(vt);
    }
}

/* LSD radix sort with a stable scatter (radix_sort.hpp:186-250, rank =
 * earlier same-digit items, :234-239) orders the keys stably by the full
 * transformed key radix(x); this is that order, computed by comparison. */
void orc_radix_order_sort(void *keys, void *vals, size_t n, int kbytes,
                          int kind, int descending, int vbytes)
{
    if (n < 2) return;
    uint64_t *rk = (uint64_t *)malloc(n * sizeof(uint64_t));
    size_t *idx = (size_t *)malloc(n * sizeof(size_t));
    for (size_t i = 0; i < n; ++i) {
        rk[i] = orc_radix_key(load_bits(keys, i, kbytes), kbytes, kind, descending);
        idx[i] = i;
    }
    stable_index_sort(rk, idx, n);
    apply_perm(keys, vals, idx, n, kbytes, vbytes);
    free(rk);
    free(idx);
}

/* native "a < b" for the key type described by (kbytes, kind) */
static int native_less(uint64_t a, uint64_t b, int kbytes, int kind)
{
    if (kind == 2) {
        if (kbytes == 4) {
            float fa, fb;
            uint32_t ua = (uint32_t)a, ub = (uint32_t)b;
            memcpy(&fa, &ua, 4);
            memcpy(&fb, &ub, 4);
            return fa < fb;
        } else {
            double da, db;
            memcpy(&da, &a, 8);
            memcpy(&db, &b, 8);
            return da < db;
        }
    }
    if (kind == 1) {
        int W = kbytes * 8;
        int64_t sa = (W == 64) ? (int64_t)a : (int64_t)(a << (64 - W)) >> (64 - W);
        int64_t sb = (W == 64) ? (int64_t)b : (int64_t)(b << (64 - W)) >> (64 - W);
        return sa < sb;
    }
    return a < b;
}

static int native_cmp(uint64_t v, uint64_t ref, int kbytes, int kind, int descending)
{
    /* compare(value, data[pos-1]): less<T> -> value < ref; greater<T> -> value > ref */
    return descending ? native_less(ref, v, kbytes, kind) : native_less(v, ref, kbytes, kind);
}

/* serial_insertion_sort(_by_key), algorithm/detail/insertion_sort.hpp:42-64
 * and :106-133: shift while compare(value, data[pos-1]). */
void orc_insertion_sort(void *keys, void *vals, size_t n, int kbytes, int kind,
                        int descending, int vbytes)
{
    if (n < 2) return;
    char vbuf[16];
    for (size_t i = 1; i < n; ++i) {
        uint64_t key = load_bits(keys, i, kbytes);
        if (vals && vbytes) memcpy(vbuf, (char *)vals + i * vbytes, (size_t)vbytes);
        size_t pos = i;
        while (pos > 0 && native_cmp(key, load_bits(keys, pos - 1, kbytes), kbytes, kind, descending)) {
            store_bits(keys, pos, kbytes, load_bits(keys, pos - 1, kbytes));
            if (vals && vbytes)
                memcpy((char *)vals + pos * vbytes, (char *)vals + (pos - 1) * vbytes, (size_t)vbytes);
            pos--;
        }
        store_bits(keys, pos, kbytes, key);
        if (vals && vbytes) memcpy((char *)vals + pos * vbytes, vbuf, (size_t)vbytes);
    }
}

/* GPU-column dispatch of the public sorts:
 *   sort:               n<2 nothing, n<=32 insertion, else radix  (sort.hpp:34-81)
 *   sort_by_key:        n<32 insertion, else radix                (sort_by_key.hpp:33-86)
 *   stable_sort:        radix at any n                            (stable_sort.hpp:52-71)
 *   stable_sort_by_key: n<32 insertion, else radix                (stable_sort_by_key.hpp:29-77) */
void orc_sort_dispatch(int api, void *keys, void *vals, size_t n, int kbytes,
                       int kind, int descending, int vbytes)
{
    int insertion = 0;
    if (api == 0) insertion = (n <= 32);
    else if (api == 1 || api == 3) insertion = (n < 32);
    if (insertion) orc_insertion_sort(keys, vals, n, kbytes, kind, descending, vbytes);
    else orc_radix_order_sort(keys, vals, n, kbytes, kind, descending, vbytes);
}

/* CPU-device column: stable comparison merge sort with operator< / operator>
 * (merge_sort_on_cpu.hpp:236-360).  Exact only for NaN-free input. */
void orc_compare_stable_sort(void *keys, void *vals, size_t n, int kbytes,
                             int kind, int descending, int vbytes)
{
    if (n < 2) return;
    size_t *idx = (size_t *)malloc(n * sizeof(size_t));
    size_t *tmp = (size_t *)malloc(n * sizeof(size_t));
    uint64_t *kb = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) { idx[i] = i; kb[i] = load_bits(keys, i, kbytes); }
    size_t *a = idx, *b = tmp;
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n;
            size_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                if (native_cmp(kb[a[j]], kb[a[i]], kbytes, kind, descending)) b[o++] = a[j++];
                else b[o++] = a[i++];
            }
            while (i < mid) b[o++] = a[i++];
            while (j < hi) b[o++] = a[j++];
        }
        size_t *t = a; a = b; b = t;
    }
    apply_perm(keys, vals, a, n, kbytes, vbytes);
    free(idx);
    free(tmp);
    free(kb);
}

/* Segmented sort: each segment as one stable_sort_by_key call would leave
 * it (the per-call baseline: stable_sort_by_key.hpp:29-77 per segment). */
void orc_segmented_sort(void *keys, void *vals, const uint64_t *offsets,
                        size_t nseg, int kbytes, int kind, int descending,
                        int vbytes)
{
    for (size_t s = 0; s < nseg; ++s) {
        size_t b = (size_t)offsets[s], e = (size_t)offsets[s + 1];
        orc_sort_dispatch(3, (char *)keys + b * kbytes,
                          vals ? (char *)vals + b * vbytes : NULL, e - b,
                          kbytes, kind, descending, vbytes);
    }
}

/* ======================================================================
 * Element arithmetic for scans / reductions / transforms
 * ==================================================================== */
static int dt_size(int dt)
{
    switch (dt) {
    case ORC_I8: case ORC_U8: return 1;
    case ORC_I16: case ORC_U16: return 2;
    case ORC_I32: case ORC_U32: case ORC_F32: return 4;
    default: return 8;
    }
}

typedef union { int64_t i; uint64_t u; double d; float f; } num_t;

static num_t ld(int dt, const void *p, size_t i)
{
    num_t v; v.u = 0;
    switch (dt) {
    case ORC_I8: v.i = ((const int8_t *)p)[i]; break;
    case ORC_U8: v.u = ((const uint8_t *)p)[i]; break;
    case ORC_I16: v.i = ((const int16_t *)p)[i]; break;
    case ORC_U16: v.u = ((const uint16_t *)p)[i]; break;
    case ORC_I32: v.i = ((const int32_t *)p)[i]; break;
    case ORC_U32: v.u = ((const uint32_t *)p)[i]; break;
    case ORC_I64: v.i = ((const int64_t *)p)[i]; break;
    case ORC_U64: v.u = ((const uint64_t *)p)[i]; break;
    case ORC_F32: v.f = ((const float *)p)[i]; break;
    default: v.d = ((const double *)p)[i]; break;
    }
    return v;
}

static void st(int dt, void *p, size_t i, num_t v)
{
    switch (dt) {
    case ORC_I8: ((int8_t *)p)[i] = (int8_t)v.i; break;
    case ORC_U8: ((uint8_t *)p)[i] = (uint8_t)v.u; break;
    case ORC_I16: ((int16_t *)p)[i] = (int16_t)v.i; break;
    case ORC_U16: ((uint16_t *)p)[i] = (uint16_t)v.u; break;
    case ORC_I32: ((int32_t *)p)[i] = (int32_t)(uint32_t)v.u; break;
    case ORC_U32: ((uint32_t *)p)[i] = (uint32_t)v.u; break;
    case ORC_I64: ((int64_t *)p)[i] = v.i; break;
    case ORC_U64: ((uint64_t *)p)[i] = v.u; break;
    case ORC_F32: ((float *)p)[i] = v.f; break;
    default: ((double *)p)[i] = v.d; break;
    }
}

static int is_signed_dt(int dt) { return dt == ORC_I8 || dt == ORC_I16 || dt == ORC_I32 || dt == ORC_I64; }

/* a op b in the element type: integers wrap modulo 2^W (OpenCL C integer
 * semantics); floats round once per operation. */
static num_t apply(int dt, int op, num_t a, num_t b)
{
    num_t r; r.u = 0;
    if (dt == ORC_F32) {
        switch (op) {
        case ORC_PLUS: r.f = a.f + b.f; break;
        case ORC_MULTIPLIES: r.f = a.f * b.f; break;
        case ORC_MIN: r.f = (b.f < a.f) ? b.f : a.f; break;
        default: r.f = (a.f < b.f) ? b.f : a.f; break;
        }
        return r;
    }
    if (dt == ORC_F64) {
        switch (op) {
        case ORC_PLUS: r.d = a.d + b.d; break;
        case ORC_MULTIPLIES: r.d = a.d * b.d; break;
        case ORC_MIN: r.d = (b.d < a.d) ? b.d : a.d; break;
        default: r.d = (a.d < b.d) ? b.d : a.d; break;
        }
        return r;
    }
    int W = dt_size(dt) * 8;
    uint64_t mask = (W == 64) ? ~0ull : ((1ull << W) - 1);
    switch (op) {
    case ORC_PLUS: r.u = (a.u + b.u) & mask; break;
    case ORC_MULTIPLIES: r.u = (a.u * b.u) & mask; break;
    case ORC_MIN:
        if (is_signed_dt(dt)) r = (b.i < a.i) ? b : a;
        else r = ((b.u & mask) < (a.u & mask)) ? b : a;
        break;
    default:
        if (is_signed_dt(dt)) r = (a.i < b.i) ? b : a;
        else r = ((a.u & mask) < (b.u & mask)) ? b : a;
        break;
    }
    if (is_signed_dt(dt) && W < 64) {
        /* re-sign-extend the wrapped value */
        r.i = (int64_t)(r.u << (64 - W)) >> (64 - W);
    }
    return r;
}

/* ======================================================================
 * Scans.  Sequential left folds: for integers every association gives the
 * same bits, so this is the reference's result (exclusive_scan.hpp:55-104,
 * inclusive_scan.hpp:53-87; exclusive element i = init op x0 op ... op
 * x(i-1), inclusive = x0 op ... op xi).  The reference's GPU kernel seeds
 * every block with 0 (scan_on_gpu.hpp:61-65), which is the identity only for
 * plus; this oracle uses the mathematical definition for every op.
 * ==================================================================== */
void orc_scan(int dtype, int op, const void *in, void *out, size_t n,
              const void *init, int exclusive)
{
    if (n == 0) return;
    num_t acc;
    if (exclusive) {
        acc = ld(dtype, init, 0);
        for (size_t i = 0; i < n; ++i) {
            num_t x = ld(dtype, in, i); /* read before write: in-place safe */
            st(dtype, out, i, acc);
            acc = apply(dtype, op, acc, x);
        }
    } else {
        acc = ld(dtype, in, 0);
        st(dtype, out, 0, acc);
        for (size_t i = 1; i < n; ++i) {
            acc = apply(dtype, op, acc, ld(dtype, in, i));
            st(dtype, out, i, acc);
        }
    }
}

/* error-free transformation: s + e == a + b exactly (Knuth's TwoSum) */
static inline void two_sum(double a, double b, double *s, double *e)
{
    double t = a + b;
    double bp = t - a;
    *e = (a - (t - bp)) + (b - bp);
    *s = t;
}

/* double-double accumulate (hi, lo) += x; no reference counterpart: this is
 * the FP64 oracle the float results are graded against. */
static inline void dd_add(double *hi, double *lo, double x)
{
    double s, e;
    two_sum(*hi, x, &s, &e);
    e += *lo;
    two_sum(s, e, hi, lo);
}

void orc_scan_f32_dd(const float *in, double *out, size_t n, int exclusive, double init)
{
    double hi = exclusive ? init : 0.0, lo = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double x = (double)in[i];
        if (exclusive) { out[i] = hi + lo; dd_add(&hi, &lo, x); }
        else { dd_add(&hi, &lo, x); out[i] = hi + lo; }
    }
}

static int keq_i32(const int32_t *k, size_t a, size_t b) { return k[a] == k[b]; }

/* Segmented scan over runs of adjacent-equal int32 keys.  No reference
 * primitive: the reference's only segmented building block is the
 * carry-out/carry-in structure of reduce_by_key_with_scan.hpp:104-293.
 * Exclusive: each run starts at init; inclusive: running fold in the run. */
void orc_segmented_scan_i32keys(const int32_t *keys, int vdtype, const void *vals,
                                void *out, size_t n, int exclusive, const void *init)
{
    num_t acc; acc.u = 0;
    for (size_t i = 0; i < n; ++i) {
        int head = (i == 0) || !keq_i32(keys, i, i - 1);
        num_t x = ld(vdtype, vals, i);
        if (exclusive) {
            if (head) acc = ld(vdtype, init, 0);
            st(vdtype, out, i, acc);
            acc = apply(vdtype, ORC_PLUS, acc, x);
        } else {
            acc = head ? x : apply(vdtype, ORC_PLUS, acc, x);
            st(vdtype, out, i, acc);
        }
    }
}

void orc_segmented_scan_f32_dd(const int32_t *keys, const float *vals, double *out,
                               size_t n, int exclusive)
{
    double hi = 0.0, lo = 0.0;
    for (size_t i = 0; i < n; ++i) {
        int head = (i == 0) || keys[i] != keys[i - 1];
        if (head) { hi = 0.0; lo = 0.0; }
        if (exclusive) { out[i] = hi + lo; dd_add(&hi, &lo, (double)vals[i]); }
        else { dd_add(&hi, &lo, (double)vals[i]); out[i] = hi + lo; }
    }
}

/* ======================================================================
 * Reductions
 * ==================================================================== */
/* Left fold (serial_reduce.hpp:42-56: result = x0; result = f(result, xi)). */
void orc_reduce(int dtype, int op, const void *in, size_t n, void *out)
{
    if (n == 0) return;
    num_t acc = ld(dtype, in, 0);
    for (size_t i = 1; i < n; ++i) acc = apply(dtype, op, acc, ld(dtype, in, i));
    st(dtype, out, 0, acc);
}

/* Double-double sum (the FP64 oracle).  Returns hi; *lo gets the tail. */
double orc_sum_f32_dd(const float *x, size_t n, double *lo)
{
    double hi = 0.0, l = 0.0;
    for (size_t i = 0; i < n; ++i) dd_add(&hi, &l, (double)x[i]);
    if (lo) *lo = l;
    return hi;
}

double orc_sum_f64_dd(const double *x, size_t n, double *lo)
{
    double hi = 0.0, l = 0.0;
    for (size_t i = 0; i < n; ++i) dd_add(&hi, &l, x[i]);
    if (lo) *lo = l;
    return hi;
}

/* reduce_on_gpu association for plus<float> (reduce_on_gpu.hpp:169-280):
 * VPT=8, TPB=128 (:220-221); per work-item strided fold from 0
 * (:186-193), then the pairwise LDS tree of ReduceBody<T,false> (:30-47);
 * passes until count <= VPT*TPB, then one final group (:262-279).  The
 * work-size arithmetic in float (:243,247,268) is exact for the sizes used. */
static size_t gpu_reduce_pass(const float *in, size_t count, float *out)
{
    const size_t VPT = 8, TPB = 128;
    size_t groups = (count + VPT * TPB - 1) / (VPT * TPB);
    float scratch[128];
    for (size_t g = 0; g < groups; ++g) {
        size_t off = g * VPT * TPB;
        for (size_t lid = 0; lid < TPB; ++lid) {
            float sum = 0.0f;
            for (size_t i = 0; i < VPT; ++i)
                if (off + lid + i * TPB < count) sum = sum + in[off + lid + i * TPB];
            scratch[lid] = sum;
        }
        for (size_t i = 1; i < TPB; i <<= 1) {
            size_t mask = (i << 1) - 1;
            for (size_t lid = 0; lid < TPB; ++lid)
                if ((lid & mask) == 0) scratch[lid] += scratch[lid + i];
        }
        out[g] = scratch[0];
    }
    return groups;
}

float orc_reduce_gpu_order_f32(const float *x, size_t n)
{
    if (n == 0) return 0.0f;
    const size_t B = 8 * 128;
    size_t cap = (n + B - 1) / B + 1;
    float *ping = (float *)malloc(cap * sizeof(float));
    float *pong = (float *)malloc(cap * sizeof(float));
    size_t count = gpu_reduce_pass(x, n, ping);
    while (count > B) {
        count = gpu_reduce_pass(ping, count, pong);
        float *t = ping; ping = pong; pong = t;
    }
    float r;
    gpu_reduce_pass(ping, count, &r); /* final single-group pass */
    free(ping);
    free(pong);
    return r;
}

/* reduce_on_cpu association (reduce_on_cpu.hpp:29-104): below
 * max(16384*sizeof(T), cu) a serial fold; otherwise cu chunks of
 * (uint)ceil((float)count/cu), each folded from its first element, then a
 * recursive reduce of the cu partial results. */
float orc_reduce_cpu_order_f32(const float *x, size_t n, unsigned cu)
{
    size_t thr = 16384 * sizeof(float);
    if (thr < cu) thr = cu;
    if (n < thr) {
        float r = x[0];
        for (size_t i = 1; i < n; ++i) r = r + x[i];
        return r;
    }
    float *part = (float *)malloc(cu * sizeof(float));
    unsigned block = (unsigned)ceilf((float)n / (float)cu);
    for (unsigned g = 0; g < cu; ++g) {
        size_t idx = (size_t)g * block;
        size_t end = idx + block < n ? idx + block : n;
        float r = idx < n ? x[idx] : 0.0f;
        for (size_t i = idx + 1; i < end; ++i) r = r + x[i];
        part[g] = r;
    }
    float r = orc_reduce_cpu_order_f32(part, cu, cu);
    free(part);
    return r;
}

/* scan_on_cpu (scan_on_cpu.hpp:28-199), the CPU-device column of the
 * scans, for plus on float (ORC_F32) or int32 (ORC_I32, wrapping).  G = cu
 * work-items; below max(16384 * sizeof(T), cu) elements it is serial_scan
 * (serial_scan.hpp:50-80): one left fold, init first when exclusive.
 * Otherwise block = (n + G) / (G + 1) (:76) and chunk c = [c*block,
 * min(n, (c+1)*block)), c = 0..G:
 *   kernel 1 (:75-124), work-item g < G, chunk g: a left fold of the chunk
 *     from its first element (exclusive, g == 0: from init instead, :86-94);
 *     only g == 0 writes its chunk's results (:104-118); every g writes its
 *     fold to partial[g] (:124) and, inclusive, result[g*block] = x[g*block]
 *     (:82-84), overwritten by kernel 2 for g >= 1;
 *   kernel 2 (:147-183), work-item g, chunk g+1: seed = partial[0] op
 *     partial[1] ... op partial[g] (a left fold, :153-157), then a left fold
 *     of the chunk from the seed (exclusive writes the seed first, :161-170).
 *   With G < 2 kernel 2 never runs (:135-137): the chunk after the first is
 *   left unwritten, as in the reference.
 * A work-item of kernel 1 whose chunk is empty returns before writing its
 * partial (:79), which a later seed would read uninitialised; that needs
 * n <= G(G-1)/2 and returns -1 here (results undefined in the reference).
 * Returns 0 otherwise. */
#define ORC_SCAN_CPU(NAME, T, ADD)                                                          \
static int NAME(const T *x, T *out, size_t n, int exclusive, T init, unsigned G)            \
{                                                                                           \
    size_t thr = 16384 * sizeof(T);                                                         \
    if (thr < G) thr = G;                                                                   \
    if (n == 0) return 0;                                                                   \
    if (n < thr) {                                                                          \
        T sum = exclusive ? init : x[0];                                                    \
        size_t i = 1;                                                                       \
        if (exclusive) i = 0; else out[0] = sum;                                            \
        for (; i < n; ++i) {                                                                \
            const T v = x[i];                                                               \
            if (exclusive) out[i] = sum;                                                    \
            sum = ADD(sum, v);                                                              \
            if (!exclusive) out[i] = sum;                                                   \
        }                                                                                   \
        return 0;                                                                           \
    }                                                                                       \
    const uint32_t cnt = (uint32_t)n;                                                       \
    const uint32_t block = (cnt + G) / (G + 1);                                             \
    T *part = (T *)malloc(G * sizeof(T));                                                   \
    int undefined = 0;                                                                      \
    for (uint32_t g = 0; g < G; ++g) {                                                      \
        uint32_t index = g * block;                                                         \
        const uint32_t end = cnt < index + block ? cnt : index + block;                     \
        if (index >= end) { undefined = 1; continue; }                                      \
        T sum;                                                                              \
        if (!exclusive) {                                                                   \
            sum = x[index];                                                                 \
            out[index] = sum;                                                               \
            index++;                                                                        \
        } else if (index == 0) {                                                            \
            sum = init;                                                                     \
        } else {                                                                            \
            sum = x[index];                                                                 \
            index++;                                                                        \
        }                                                                                   \
        for (; index < end; ++index) {                                                      \
            const T v = x[index];                                                           \
            if (exclusive && g == 0) out[index] = sum;                                      \
            sum = ADD(sum, v);                                                              \
            if (!exclusive && g == 0) out[index] = sum;                                     \
        }                                                                                   \
        part[g] = sum;                                                                      \
    }                                                                                       \
    if (G >= 2) {                                                                           \
        for (uint32_t g = 0; g < G; ++g) {                                                  \
            uint32_t index = block + g * block;                                             \
            const uint32_t end = cnt < index + block ? cnt : index + block;                 \
            T sum = part[0];                                                                \
            for (uint32_t i = 0; i < g; ++i) sum = ADD(sum, part[i + 1]);                   \
            for (; index < end; ++index) {                                                  \
                if (exclusive) {                                                            \
                    const T v = x[index];                                                   \
                    out[index] = sum;                                                       \
                    sum = ADD(sum, v);                                                      \
                } else {                                                                    \
                    sum = ADD(sum, x[index]);                                               \
                    out[index] = sum;                                                       \
                }                                                                           \
            }                                                                               \
        }                                                                                   \
    }                                                                                       \
    free(part);                                                                             \
    return undefined ? -1 : 0;                                                              \
}

static float orc_addf(float a, float b) { return a + b; }
static int32_t orc_addi(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
ORC_SCAN_CPU(scan_cpu_f32, float, orc_addf)
ORC_SCAN_CPU(scan_cpu_i32, int32_t, orc_addi)
#undef ORC_SCAN_CPU

int orc_scan_cpu_order(int dtype, const void *in, void *out, size_t n, int exclusive, const void *init,
                       unsigned cu)
{
    if (dtype == ORC_F32)
        return scan_cpu_f32((const float *)in, (float *)out, n, exclusive, init ? *(const float *)init : 0.0f, cu);
    if (dtype == ORC_I32)
        return scan_cpu_i32((const int32_t *)in, (int32_t *)out, n, exclusive, init ? *(const int32_t *)init : 0,
                            cu);
    return -2;
}

/* serial_accumulate (serial_accumulate.hpp:38-40): result = init; result += xi */
float orc_accumulate_f32(const float *x, size_t n, float init)
{
    float r = init;
    for (size_t i = 0; i < n; ++i) r = r + x[i];
    return r;
}

/* scan_on_gpu association for plus<float> (scan_on_gpu.hpp:187-267):
 * blocks of pick_scan_block_size(n) (<= 256) scanned Hillis-Steele in LDS
 * (:84-93); block sums scanned recursively (inclusive, :240-249, in-place
 * via a copy :281-303) and added to every later block (:132-168). */
static size_t pick_bs(size_t n)
{
    size_t b = 1;
    while (b < n && b < 256) b <<= 1;
    return n == 0 ? 0 : b;
}

static void gpu_scan_impl(const float *in, float *out, size_t count, int exclusive, float init)
{
    size_t bs = pick_bs(count);
    size_t nb = (count + bs - 1) / bs;
    float *sums = (float *)calloc(nb, sizeof(float));
    float *scr = (float *)malloc(bs * sizeof(float));
    float *snap = (float *)malloc(bs * sizeof(float));
    for (size_t b = 0; b < nb; ++b) {
        for (size_t lid = 0; lid < bs; ++lid) {
            size_t gid = b * bs + lid;
            if (gid < count) {
                if (exclusive) scr[lid] = (lid == 0) ? (gid == 0 ? init : 0.0f) : in[gid - 1];
                else scr[lid] = in[gid];
            } else {
                scr[lid] = 0.0f;
            }
        }
        for (size_t i = 1; i < bs; i <<= 1) {
            for (size_t lid = 0; lid < bs; ++lid) snap[lid] = lid >= i ? scr[lid - i] : 0.0f;
            for (size_t lid = 0; lid < bs; ++lid) if (lid >= i) scr[lid] = scr[lid] + snap[lid];
        }
        for (size_t lid = 0; lid < bs; ++lid) {
            size_t gid = b * bs + lid;
            if (gid < count) out[gid] = scr[lid];
        }
        size_t last = b * bs + bs - 1;
        if (exclusive) {
            if (last < count) sums[b] = in[last] + scr[bs - 1];
        } else {
            sums[b] = scr[bs - 1];
        }
    }
    if (nb > 1) {
        float *tmp = (float *)malloc(nb * sizeof(float));
        memcpy(tmp, sums, nb * sizeof(float));
        gpu_scan_impl(tmp, sums, nb, 0, init);
        free(tmp);
        for (size_t gid = bs; gid < count; ++gid) out[gid] = sums[gid / bs - 1] + out[gid];
    }
    free(sums);
    free(scr);
    free(snap);
}

void orc_scan_gpu_order_f32(const float *x, float *out, size_t n, int exclusive)
{
    if (n == 0) return;
    gpu_scan_impl(x, out, n, exclusive, 0.0f);
}

float orc_scan_gpu_order_f32_last(const float *x, size_t n)
{
    float *o = (float *)malloc(n * sizeof(float));
    orc_scan_gpu_order_f32(x, o, n, 0);
    float r = o[n - 1];
    free(o);
    return r;
}

/* ======================================================================
 * reduce_by_key: serial_reduce_by_key (serial_reduce_by_key.hpp:53-97):
 * fold each run of predicate-equal adjacent keys left to right; emit the
 * run's LAST key (previous_key at the boundary, :80,87) and its fold.
 * equal_to on floats: -0 == +0, NaN never equal.
 * ==================================================================== */
static int key_equal(int kdt, num_t a, num_t b)
{
    if (kdt == ORC_F32) return a.f == b.f;
    if (kdt == ORC_F64) return a.d == b.d;
    int W = dt_size(kdt) * 8;
    uint64_t mask = (W == 64) ? ~0ull : ((1ull << W) - 1);
    return (a.u & mask) == (b.u & mask);
}

size_t orc_reduce_by_key(int kdtype, int vdtype, int op, const void *keys,
                         const void *vals, size_t n, void *keys_out, void *vals_out)
{
    if (n == 0) return 0;
    num_t result = ld(vdtype, vals, 0);
    num_t prev = ld(kdtype, keys, 0);
    size_t size = 1;
    for (size_t i = 1; i < n; ++i) {
        num_t v = ld(vdtype, vals, i);
        num_t k = ld(kdtype, keys, i);
        if (key_equal(kdtype, prev, k)) {
            result = apply(vdtype, op, result, v);
        } else {
            st(kdtype, keys_out, size - 1, prev);
            st(vdtype, vals_out, size - 1, result);
            result = v;
            size++;
        }
        prev = k;
    }
    st(kdtype, keys_out, size - 1, prev);
    st(vdtype, vals_out, size - 1, result);
    return size;
}

/* per-run double-double sums (FP64 oracle for float reduce_by_key) */
size_t orc_reduce_by_key_f32_dd(const int32_t *keys, const float *vals, size_t n,
                                int32_t *keys_out, double *sums)
{
    if (n == 0) return 0;
    size_t r = 0;
    double hi = 0.0, lo = 0.0;
    for (size_t i = 0; i < n; ++i) {
        if (i > 0 && keys[i] != keys[i - 1]) {
            keys_out[r] = keys[i - 1];
            sums[r] = hi + lo;
            r++;
            hi = 0.0; lo = 0.0;
        }
        dd_add(&hi, &lo, (double)vals[i]);
    }
    keys_out[r] = keys[n - 1];
    sums[r] = hi + lo;
    return r + 1;
}

/* ======================================================================
 * transform.  saxpy is the reference's lambda a*_1+_2 printed as one
 * OpenCL C expression "(a*x[i])+y[i]" (lambda/context.hpp:34-79,
 * perf/perf_saxpy.cpp:50-55); OpenCL's default FP_CONTRACT ON lets the
 * compiler fuse it, and this oracle fixes the fused form fmaf(a, x, y).
 * ==================================================================== */
void orc_saxpy_f32(float a, const float *x, const float *y, float *out, size_t n)
{
    for (size_t i = 0; i < n; ++i) out[i] = fmaf(a, x[i], y[i]);
}

void orc_transform(int dtype, int op, const void *a, const void *b, void *out, size_t n)
{
    for (size_t i = 0; i < n; ++i) st(dtype, out, i, apply(dtype, op, ld(dtype, a, i), ld(dtype, b, i)));
}

/* ======================================================================
 * Selection.  The reference prints a lambda predicate as an OpenCL C
 * expression (lambda/context.hpp:145-156) that the device compiler
 * evaluates with C's usual arithmetic conversions; this restates that
 * evaluation for the recognised forms f(_1) cmp c.
 * ==================================================================== */
static int promoted_i32(int dt) { return dt == ORC_I8 || dt == ORC_U8 || dt == ORC_I16 || dt == ORC_U16; }

int orc_pred_eval(int dtype, const void *in, size_t i, const orc_predicate *p)
{
    num_t x = ld(dtype, in, i);
    int lt, eq, gt, r;
    if (p->fn == ORC_FN_SIGNBIT) {
        /* signbit (functional/relational.hpp:32): OpenCL's int result */
        int32_t s = dtype == ORC_F32 ? (int32_t)signbit(x.f) != 0 : (int32_t)signbit(x.d) != 0;
        int32_t c; memcpy(&c, &p->value, 4);
        lt = s < c; eq = s == c; gt = s > c;
    } else if (promoted_i32(dtype) || dtype == ORC_I32) {
        int32_t y = promoted_i32(dtype) ? (int32_t)(is_signed_dt(dtype) ? x.i : (int64_t)x.u) : (int32_t)x.i;
        int32_t a, c;
        memcpy(&a, &p->operand, 4);
        memcpy(&c, &p->value, 4);
        if (p->fn == ORC_FN_MUL) y = (int32_t)((uint32_t)y * (uint32_t)a);
        else if (p->fn == ORC_FN_MOD) y = (a == -1) ? 0 : y % a;
        else if (p->fn == ORC_FN_AND) y = y & a;
        lt = y < c; eq = y == c; gt = y > c;
    } else if (dtype == ORC_U32) {
        uint32_t y = (uint32_t)x.u, a, c;
        memcpy(&a, &p->operand, 4);
        memcpy(&c, &p->value, 4);
        if (p->fn == ORC_FN_MUL) y = y * a;
        else if (p->fn == ORC_FN_MOD) y = y % a;
        else if (p->fn == ORC_FN_AND) y = y & a;
        lt = y < c; eq = y == c; gt = y > c;
    } else if (dtype == ORC_I64) {
        int64_t y = x.i, a, c;
        memcpy(&a, &p->operand, 8);
        memcpy(&c, &p->value, 8);
        if (p->fn == ORC_FN_MUL) y = (int64_t)((uint64_t)y * (uint64_t)a);
        else if (p->fn == ORC_FN_MOD) y = (a == -1) ? 0 : y % a;
        else if (p->fn == ORC_FN_AND) y = y & a;
        lt = y < c; eq = y == c; gt = y > c;
    } else if (dtype == ORC_U64) {
        uint64_t y = x.u, a = p->operand, c = p->value;
        if (p->fn == ORC_FN_MUL) y = y * a;
        else if (p->fn == ORC_FN_MOD) y = y % a;
        else if (p->fn == ORC_FN_AND) y = y & a;
        lt = y < c; eq = y == c; gt = y > c;
    } else if (dtype == ORC_F32) {
        float y = x.f, a, c;
        memcpy(&a, &p->operand, 4);
        memcpy(&c, &p->value, 4);
        if (p->fn == ORC_FN_MUL) y = y * a;
        lt = y < c; eq = y == c; gt = y > c;
    } else {
        double y = x.d, a, c;
        memcpy(&a, &p->operand, 8);
        memcpy(&c, &p->value, 8);
        if (p->fn == ORC_FN_MUL) y = y * a;
        lt = y < c; eq = y == c; gt = y > c;
    }
    switch (p->cmp) {
    case ORC_CMP_LT: r = lt; break;
    case ORC_CMP_LE: r = lt || eq; break;
    case ORC_CMP_GT: r = gt; break;
    case ORC_CMP_GE: r = gt || eq; break;
    case ORC_CMP_EQ: r = eq; break;
    default: r = !eq; break;
    }
    return p->negate ? !r : r;
}

/* transform_if_impl (transform_if.hpp:33-82) with the identity function:
 * k1 writes indices[i] = pred(first[i]) ? 1 : 0 (:54-57); the count is the
 * last flag plus the last element of the exclusive scan of the flags
 * (:60-64); k2 writes result[indices[i]] = first[i], or i for
 * copy_index_if (copy_if.hpp:25-41), where pred holds (:67-78). */
size_t orc_copy_if(int dtype, const void *in, size_t n, const orc_predicate *p, int out_kind, void *out)
{
    if (n == 0) return 0;
    uint64_t *idx = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) idx[i] = orc_pred_eval(dtype, in, i, p) ? 1 : 0;
    uint64_t copied = idx[n - 1], acc = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t f = idx[i];
        idx[i] = acc;
        acc += f;
    }
    copied += idx[n - 1];
    for (size_t i = 0; i < n; ++i) {
        if (!orc_pred_eval(dtype, in, i, p)) continue;
        if (out_kind == 1) ((uint32_t *)out)[idx[i]] = (uint32_t)i;
        else if (out_kind == 2) ((uint64_t *)out)[idx[i]] = (uint64_t)i;
        else st(dtype, out, idx[i], ld(dtype, in, i));
    }
    free(idx);
    return (size_t)copied;
}

/* count_if (count_if.hpp; count_if_with_reduce.hpp sums pred(x) as 0/1) */
size_t orc_count_if(int dtype, const void *in, size_t n, const orc_predicate *p)
{
    size_t c = 0;
    for (size_t i = 0; i < n; ++i) c += orc_pred_eval(dtype, in, i, p) ? 1 : 0;
    return c;
}

/* unique_copy with equal_to (unique_copy.hpp:139-164): below 32 elements
 * serial_unique_copy (:26-73) keeps next when !(current == next) and then
 * makes it current; otherwise flags[0] = 1, flags[i] = !(x[i-1] == x[i])
 * (:114-120) and the flagged elements are gathered in order (:122-131). */
size_t orc_unique_copy(int dtype, const void *in, size_t n, void *out)
{
    if (n == 0) return 0;
    if (n < 32) {
        size_t index = 0;
        num_t current = ld(dtype, in, 0);
        st(dtype, out, 0, current);
        for (size_t i = 1; i < n; ++i) {
            num_t next = ld(dtype, in, i);
            if (!key_equal(dtype, current, next)) {
                st(dtype, out, ++index, next);
                current = next;
            }
        }
        return index + 1;
    }
    size_t r = 0;
    for (size_t i = 0; i < n; ++i) {
        int flag = i == 0 || !key_equal(dtype, ld(dtype, in, i - 1), ld(dtype, in, i));
        if (flag) st(dtype, out, r++, ld(dtype, in, i));
    }
    return r;
}

/* partition_copy (partition_copy.hpp): copy_if(pred) to first_true, then
 * copy_if(not1(pred)) to first_false. */
size_t orc_partition_copy(int dtype, const void *in, size_t n, const orc_predicate *p, void *out_true,
                          void *out_false)
{
    orc_predicate q = *p;
    q.negate = !p->negate;
    size_t t = orc_copy_if(dtype, in, n, p, 0, out_true);
    orc_copy_if(dtype, in, n, &q, 0, out_false);
    return t;
}

/* stable_partition (stable_partition.hpp): copy the range to a temporary,
 * copy_if(pred) from it to first, then copy_if(not1(pred)) after the last
 * true; returns the number of true elements.  partition (partition.hpp)
 * is the same call. */
size_t orc_stable_partition(int dtype, void *data, size_t n, const orc_predicate *p)
{
    if (n == 0) return 0;
    size_t bytes = n * (size_t)dt_size(dtype);
    void *tmp = malloc(bytes);
    memcpy(tmp, data, bytes);
    orc_predicate q = *p;
    q.negate = !p->negate;
    size_t t = orc_copy_if(dtype, tmp, n, p, 0, data);
    orc_copy_if(dtype, tmp, n, &q, 0, (char *)data + t * (size_t)dt_size(dtype));
    free(tmp);
    return t;
}

/* ======================================================================
 * transform_reduce (transform_reduce.hpp: reduce over
 * make_transform_iterator(first, f), or over a zip_iterator of two ranges)
 * and inner_product (inner_product.hpp: accumulate over the products of a
 * zipped pair; its general form transforms into a temporary and
 * accumulates).  The element map is evaluated in the element type, one
 * rounding (or wrap) per operation, as the printed OpenCL C expression is.
 * counting_iterator (iterator/counting_iterator.hpp) yields init + i.
 * ==================================================================== */
static num_t map_elem(int dt, const void *p, int64_t start, size_t i)
{
    if (p) return ld(dt, p, i);
    num_t v; v.u = 0;
    if (dt == ORC_F32) v.f = (float)start + (float)i;
    else if (dt == ORC_F64) v.d = (double)start + (double)i;
    else {
        int W = dt_size(dt) * 8;
        uint64_t mask = (W == 64) ? ~0ull : ((1ull << W) - 1);
        uint64_t u = ((uint64_t)start + (uint64_t)i) & mask;
        if (is_signed_dt(dt) && W < 64 && (u >> (W - 1))) u |= ~mask;  /* sign-extend */
        v.u = u;
    }
    return v;
}

static num_t map_sub_(int dt, num_t a, num_t b)
{
    num_t r; r.u = 0;
    if (dt == ORC_F32) { r.f = a.f - b.f; return r; }
    if (dt == ORC_F64) { r.d = a.d - b.d; return r; }
    int W = dt_size(dt) * 8;
    uint64_t mask = (W == 64) ? ~0ull : ((1ull << W) - 1);
    uint64_t u = (a.u - b.u) & mask;
    if (is_signed_dt(dt) && W < 64 && (u >> (W - 1))) u |= ~mask;
    r.u = u;
    return r;
}

static num_t map_abs_(int dt, num_t a)
{
    num_t r = a;
    if (dt == ORC_F32) { r.f = fabsf(a.f); return r; }
    if (dt == ORC_F64) { r.d = fabs(a.d); return r; }
    if (is_signed_dt(dt) && a.i < 0) {
        num_t z; z.u = 0;
        return map_sub_(dt, z, a);  /* wraps: abs(INT_MIN) keeps INT_MIN's bits */
    }
    return r;
}

/* the value g is applied to: a op b, or a - c */
static num_t map_inner(int dt, num_t a, num_t b, int binary, const orc_map *m)
{
    if (!binary) return map_sub_(dt, a, ld(dt, &m->offset, 0));  /* c: one element in the low bytes */
    if (m->op == 4) return map_sub_(dt, a, b);
    return apply(dt, m->op == ORC_MULTIPLIES ? ORC_MULTIPLIES : ORC_PLUS, a, b);
}

void orc_map_reduce(int dtype, const void *a, int64_t a_start, const void *b, int64_t b_start, int binary, size_t n,
                    const orc_map *m, int rop, const void *init, void *out)
{
    int have = 0;
    num_t acc; acc.u = 0;
    if (init) { acc = ld(dtype, init, 0); have = 1; }
    for (size_t i = 0; i < n; ++i) {
        num_t x = map_elem(dtype, a, a_start, i);
        num_t y = binary ? map_elem(dtype, b, b_start, i) : x;
        num_t t;
        if (binary && m->op == ORC_MULTIPLIES && m->fn == ORC_MAP_ABS)
            t = apply(dtype, ORC_MULTIPLIES, map_abs_(dtype, x), map_abs_(dtype, y));
        else {
            t = map_inner(dtype, x, y, binary, m);
            if (m->fn == ORC_MAP_ABS) t = map_abs_(dtype, t);
            else if (m->fn == ORC_MAP_SQUARE) t = apply(dtype, ORC_MULTIPLIES, t, t);
        }
        acc = have ? apply(dtype, rop, acc, t) : t;
        have = 1;
    }
    if (have) st(dtype, out, 0, acc);
}

/* exact product of two floats of the element type as a double-double */
static void prod_dd(int dt, num_t x, num_t y, double *hi, double *lo)
{
    if (dt == ORC_F32) { *hi = (double)x.f * (double)y.f; *lo = 0.0; return; }  /* exact in double */
    *hi = x.d * y.d;
    *lo = fma(x.d, y.d, -*hi);
}

double orc_map_sum_dd(int dtype, const void *a, int64_t a_start, const void *b, int64_t b_start, int binary,
                      size_t n, const orc_map *m, double *lo_out)
{
    double hi = 0.0, lo = 0.0;
    for (size_t i = 0; i < n; ++i) {
        num_t x = map_elem(dtype, a, a_start, i);
        num_t y = binary ? map_elem(dtype, b, b_start, i) : x;
        double ph, pl;
        if (binary && m->op == ORC_MULTIPLIES && m->fn != ORC_MAP_SQUARE) {
            if (m->fn == ORC_MAP_ABS) { x = map_abs_(dtype, x); y = map_abs_(dtype, y); }
            prod_dd(dtype, x, y, &ph, &pl);
        } else {
            num_t t = map_inner(dtype, x, y, binary, m);
            if (m->fn == ORC_MAP_ABS) t = map_abs_(dtype, t);
            if (m->fn == ORC_MAP_SQUARE) prod_dd(dtype, t, t, &ph, &pl);
            else { ph = dtype == ORC_F32 ? (double)t.f : t.d; pl = 0.0; }
        }
        dd_add(&hi, &lo, ph);
        dd_add(&hi, &lo, pl);
    }
    *lo_out = lo;
    return hi;
}
