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

}  // namespace

extern "C" {

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

}  // extern "C"
