// workloads.cpp -- seeded synthetic inputs for the benchmark configurations.
//
// Host-only helper library (libbcmi_workloads.so) used by bench.py.  It uses
// the standard engines std::mt19937 / std::mt19937_64 and explicit
// transforms of their raw output (never a std::*_distribution, whose
// algorithms are implementation-defined), so every harness that follows the
// same definitions produces identical bits.  The test suite compares these
// generators with the oracle's independent implementation.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <random>

namespace {

constexpr double kTwoM32 = 2.3283064365386962890625e-10;  // 2^-32
constexpr double kTwoM53 = 1.1102230246251565404236316680908203125e-16;
constexpr double kTwoPi = 6.283185307179586476925286766559;

// Box-Muller, cosine branch: u1 in (0,1], u2 in [0,1)
double normal32(std::mt19937 &g)
{
    double u1 = ((double)g() + 1.0) * kTwoM32;
    double u2 = (double)g() * kTwoM32;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
}

double normal64(std::mt19937_64 &g)
{
    double u1 = ((double)(g() >> 11) + 1.0) * kTwoM53;
    double u2 = (double)(g() >> 11) * kTwoM53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
}

// Double-double accumulator (Knuth two-sum, then a fast renormalisation):
// about 106 bits, so its value stands in for the exact sum when the float
// results' errors are measured.
struct DD {
    double hi = 0.0, lo = 0.0;
    void add(double x)
    {
        const double s = hi + x;
        const double bp = s - hi;
        double e = (hi - (s - bp)) + (x - bp);
        e += lo;
        hi = s + e;
        lo = e - (hi - s);
    }
    double value() const { return hi + lo; }
};

// spacing of float32 at |r| (the float ulp the error is counted in)
double ulp_f32(double r)
{
    r = std::fabs(r);
    if (r < 0x1p-126) return 0x1p-149;
    int e;
    std::frexp(r, &e);  // r = m * 2^e, m in [0.5, 1)
    return std::ldexp(1.0, e - 24);
}

// Running maxima of |got - ref| (absolute, relative to |ref| where ref != 0,
// and in float ulps of ref).
struct ErrMax {
    double abs = 0.0, rel = 0.0, ulp = 0.0;
    void add(float got, double ref)
    {
        const double e = std::fabs((double)got - ref);
        if (e > abs) abs = e;
        if (ref != 0.0 && e / std::fabs(ref) > rel) rel = e / std::fabs(ref);
        const double u = e / ulp_f32(ref);
        if (u > ulp) ulp = u;
    }
    void out(double *o) const
    {
        o[0] = abs;
        o[1] = rel;
        o[2] = ulp;
    }
};

}  // namespace

extern "C" {

// ---------------------------------------------- float error measurement
// Each writes {max abs, max rel, max ulp} of a float32 result against the
// double-double value of the same operation into err[3].

// sum of x with init folded in first; returns the double-double value
double wl_sum_f32_dd(const float *x, size_t n, double init)
{
    DD a;
    a.add(init);
    for (size_t i = 0; i < n; ++i) a.add((double)x[i]);
    return a.value();
}

// double-double inclusive prefix sums of x, each rounded to double into out
void wl_scan_f32_dd(const float *x, double *out, size_t n)
{
    DD a;
    for (size_t i = 0; i < n; ++i) {
        a.add((double)x[i]);
        out[i] = a.value();
    }
}

// inclusive (exclusive != 0: exclusive, seeded with init) prefix sums
void wl_err_scan_f32(const float *x, const float *got, size_t n, int exclusive, double init, double *err)
{
    DD a;
    a.add(init);
    ErrMax m;
    for (size_t i = 0; i < n; ++i) {
        if (exclusive) {
            m.add(got[i], a.value());
            a.add((double)x[i]);
        } else {
            a.add((double)x[i]);
            m.add(got[i], a.value());
        }
    }
    m.out(err);
}

// segmented prefix sums over runs of equal keys (restarting at 0 per run)
void wl_err_segscan_f32(const int32_t *keys, const float *x, const float *got, size_t n, int exclusive, double *err)
{
    DD a;
    ErrMax m;
    for (size_t i = 0; i < n; ++i) {
        if (i > 0 && keys[i] != keys[i - 1]) a = DD();
        if (exclusive) {
            m.add(got[i], a.value());
            a.add((double)x[i]);
        } else {
            a.add((double)x[i]);
            m.add(got[i], a.value());
        }
    }
    m.out(err);
}

// per-run sums (reduce_by_key): got[r] against run r's sum; returns the
// number of runs found in keys
size_t wl_err_runsum_f32(const int32_t *keys, const float *x, size_t n, const float *got, size_t runs, double *err)
{
    ErrMax m;
    size_t r = 0;
    DD a;
    for (size_t i = 0; i < n; ++i) {
        if (i > 0 && keys[i] != keys[i - 1]) {
            if (r < runs) m.add(got[r], a.value());
            ++r;
            a = DD();
        }
        a.add((double)x[i]);
    }
    if (n) {
        if (r < runs) m.add(got[r], a.value());
        ++r;
    }
    m.out(err);
    return r;
}

// uint32 sort keys: raw mt19937(seed) output
void wl_sort_u32(uint32_t seed, uint32_t *keys, size_t n)
{
    std::mt19937 g(seed);
    for (size_t i = 0; i < n; ++i) keys[i] = (uint32_t)g();
}

// float32 keys uniform in [-1e6, 1e6]; 1% signed zeros; 1% duplicates drawn
// from the prefix; payload = index
void wl_sort_by_key_f32(float *keys, uint32_t *vals, size_t n)
{
    std::mt19937 g(0xB0057002u), g2(0xB0057012u);
    for (size_t i = 0; i < n; ++i) {
        float u = (float)(g() >> 8) * 0x1p-24f;
        float t = 2e6f * u;  // built with -ffp-contract=off: two roundings
        float k = -1e6f + t;
        uint32_t r = (uint32_t)g2();
        keys[i] = k;
        if (r % 200u == 0u) keys[i] = -0.0f;
        else if (r % 200u == 1u) keys[i] = 0.0f;
        else if (r % 100u == 2u) keys[i] = keys[(uint32_t)g2() % (uint32_t)(i + 1)];
        vals[i] = (uint32_t)i;
    }
}

// int32 uniform in [-100, 100]
void wl_scan_i32(int32_t *x, size_t n)
{
    std::mt19937 g(0xB0057003u);
    for (size_t i = 0; i < n; ++i) x[i] = (int32_t)((uint32_t)g() % 201u) - 100;
}

// x then y ~ N(0,1)
void wl_saxpy_f32(float *x, float *y, size_t n)
{
    std::mt19937 g(0xB0057004u);
    for (size_t i = 0; i < n; ++i) x[i] = (float)normal32(g);
    for (size_t i = 0; i < n; ++i) y[i] = (float)normal32(g);
}

// +-exp(3 z): lognormal(0, 3) with a random sign (low bit of the next draw)
void wl_lognormal_f32(float *x, size_t n)
{
    std::mt19937 g(0xB0057005u);
    for (size_t i = 0; i < n; ++i) {
        double v = std::exp(3.0 * normal32(g));
        if (g() & 1u) v = -v;
        x[i] = (float)v;
    }
}

// nseg lognormal segment lengths (median 256, sigma sqrt(2 ln 4)), clamped to
// [1, max_len], segment 0 = max_len; offsets[nseg+1]; returns the total
size_t wl_segments(uint64_t *offsets, size_t nseg, size_t max_len)
{
    std::mt19937_64 g(0xB0057006ull);
    const double sigma = std::sqrt(2.0 * std::log(4.0)), mu = std::log(256.0);
    offsets[0] = 0;
    for (size_t s = 0; s < nseg; ++s) {
        long long L = std::llround(std::exp(mu + sigma * normal64(g)));
        if (L < 1) L = 1;
        if ((unsigned long long)L > max_len) L = (long long)max_len;
        if (s == 0) L = (long long)max_len;
        offsets[s + 1] = offsets[s] + (uint64_t)L;
    }
    return (size_t)offsets[nseg];
}

// uint64 keys, payload = global index
void wl_segment_data(uint64_t *keys, uint32_t *vals, size_t n)
{
    std::mt19937_64 g(0xB0057016ull);
    for (size_t i = 0; i < n; ++i) {
        keys[i] = g();
        vals[i] = (uint32_t)i;
    }
}

// sorted int32 run keys (geometric lengths, p = 1/32, capped), values
// z * 10^(6u - 3); returns the number of runs
size_t wl_runs_i32_f32(int32_t *keys, float *vals, size_t n, size_t max_run)
{
    std::mt19937 g(0xB0057007u);
    const double l1p = std::log(1.0 - 1.0 / 32.0);
    size_t i = 0, run = 0;
    while (i < n) {
        double u = ((double)g() + 0.5) * kTwoM32;
        size_t L = 1 + (size_t)std::floor(std::log(u) / l1p);
        if (L > max_run) L = max_run;
        for (size_t j = 0; j < L && i < n; ++j, ++i) {
            double z = normal32(g);
            double e = 6.0 * ((double)g() * kTwoM32) - 3.0;
            keys[i] = (int32_t)run;
            vals[i] = (float)(z * std::pow(10.0, e));
        }
        ++run;
    }
    return run;
}

// ------------------------------------------------- per-call baseline loop
// One sort call per segment issued from C (the per-call baseline of the
// segmented sort, without interpreter overhead between calls).  `sort` is
// the library's bcmi_sort, passed in so this helper links nothing.
typedef int (*wl_sort_fn)(void *queue, int api, int key_dtype, void *keys, void *values, size_t value_bytes,
                          size_t n, int descending);

int wl_sort_per_segment(void *sort, void *queue, int api, int key_dtype, char *keys, size_t key_bytes, char *vals,
                        size_t value_bytes, const uint64_t *offsets, size_t nseg)
{
    wl_sort_fn fn = (wl_sort_fn)sort;
    for (size_t s = 0; s < nseg; ++s) {
        const size_t len = (size_t)(offsets[s + 1] - offsets[s]);
        if (len < 2) continue;
        const int r = fn(queue, api, key_dtype, keys + offsets[s] * key_bytes,
                         vals ? vals + offsets[s] * value_bytes : nullptr, value_bytes, len, 0);
        if (r != 0) return r;
    }
    return 0;
}

}  // extern "C"
