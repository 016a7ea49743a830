// radix_sort.hip -- sort / sort_by_key / stable_sort / stable_sort_by_key.
//
// Reference (GPU column): radix_sort_impl (algorithm/detail/radix_sort.hpp:
// 252-426): LSD radix with 4-bit digits and 128-thread blocks, i.e. 8 passes
// for 32-bit keys, each pass = count kernel + a recursive exclusive_scan of
// the per-block counts + a scan task + a scatter kernel whose rank loop is
// O(local id) (:234-239).  Below 33 (sort) or 32 (sort_by_key) elements a
// one-work-item insertion sort runs instead (sort.hpp:49-54,
// sort_by_key.hpp:48-51, insertion_sort.hpp:25-159).
//
// Here: onesweep LSD radix with 8-bit digits.  One histogram kernel reads
// the keys once and builds the digit histograms of every pass; then one
// kernel per pass: each block takes the next tile, ranks its keys with
// wave64 ballot matching against per-wave LDS digit counters, publishes its
// per-digit counts, gets its global digit offsets by decoupled look-back,
// reorders the tile in LDS so each digit run is contiguous and writes keys
// (then values) out coalesced.  The order equals the reference's: a stable
// sort by the transformed key radix(x) (radix_sort.hpp:98-129).
#include "device.hpp"

#include <cstring>
#include <type_traits>

namespace bcmi {

constexpr int kSortThreads = 256;
constexpr int kHistThreads = 256;

// radix() key transform (radix_sort.hpp:98-129) on the W-bit unsigned image
// x of the key:   key = (neg ? -x : x) ^ (flt ? (-(x >> msb) | signbit) : cx)
//   ascending:  unsigned cx=0; signed cx=signbit; float flt
//   descending: unsigned cx=~0 ((T)(-1)-x == ~x); signed neg, cx=signbit;
//               float neg, flt
template <typename KB>
struct KeyXf {
    uint32_t neg;
    uint32_t flt;
    KB cx;
};

template <typename KB>
BDEV KB radix_key(KB x, const KeyXf<KB> &f)
{
    constexpr int W = sizeof(KB) * 8;
    using SK = std::make_signed_t<KB>;
    KB t = f.neg ? (KB)(KB(0) - x) : x;
    KB sgn = (KB)(KB(1) << (W - 1));
    KB fm = f.flt ? (KB)((KB)((SK)x >> (W - 1)) | sgn) : f.cx;
    return (KB)(t ^ fm);
}

template <typename KB>
BDEV uint32_t digit_of(KB x, const KeyXf<KB> &f, uint32_t shift)
{
    return (uint32_t)(radix_key(x, f) >> shift) & 0xFFu;
}

template <int VB>
struct ValT;
template <> struct ValT<0> { using type = uint8_t; };
template <> struct ValT<1> { using type = uint8_t; };
template <> struct ValT<2> { using type = uint16_t; };
template <> struct ValT<4> { using type = uint32_t; };
template <> struct ValT<8> { using type = uint64_t; };
struct alignas(16) v16_t { uint64_t a, b; };
template <> struct ValT<16> { using type = v16_t; };

// ------------------------------------------------------------ histogram
// All passes' digit histograms in one read of the keys.  LDS counters per
// block, global atomics per block, and the last block turns the totals into
// exclusive digit offsets (ctrl->offsets) and clears the totals.
template <typename KB>
__global__ __launch_bounds__(kHistThreads) void k_histogram(const KB *__restrict__ keys, uint32_t n, KeyXf<KB> xf,
                                                            Ctrl *ctrl)
{
    constexpr int P = sizeof(KB);
    __shared__ uint32_t s_hist[P][256];
    __shared__ uint32_t s_wsum[kHistThreads / 64];
    __shared__ bool s_last;
    for (int i = threadIdx.x; i < P * 256; i += kHistThreads) (&s_hist[0][0])[i] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kHistThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kHistThreads + threadIdx.x; i < n; i += stride) {
        KB k = radix_key(keys[i], xf);
#pragma unroll
        for (int p = 0; p < P; ++p) atomicAdd(&s_hist[p][(k >> (8 * p)) & 0xFF], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P * 256; i += kHistThreads) {
        uint32_t c = (&s_hist[0][0])[i];
        if (c) __hip_atomic_fetch_add(&(&ctrl->hist[0][0])[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!arrive_last(ctrl, gridDim.x, &s_last)) return;
    const uint32_t d = threadIdx.x, lane = d & 63, w = d >> 6;
    for (int p = 0; p < P; ++p) {
        uint32_t c = __hip_atomic_load(&ctrl->hist[p][d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t inc = wave_inclusive_scan(c, 0u, [](uint32_t a, uint32_t b) { return a + b; });
        if (lane == 63) s_wsum[w] = inc;
        __syncthreads();
        uint32_t pre = 0;
        for (uint32_t i = 0; i < w; ++i) pre += s_wsum[i];
        ctrl->offsets[p][d] = pre + inc - c;
        __hip_atomic_store(&ctrl->hist[p][d], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
    if (threadIdx.x == 0) finish_launch(ctrl, false);
}

// ------------------------------------------------------------ onesweep pass
template <typename KB, int VB, int IPT>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(const KB *__restrict__ kin, KB *__restrict__ kout,
                                                           const typename ValT<VB>::type *__restrict__ vin,
                                                           typename ValT<VB>::type *__restrict__ vout, uint32_t n,
                                                           uint32_t shift, uint32_t pass, KeyXf<KB> xf, Ctrl *ctrl,
                                                           uint64_t *__restrict__ status)
{
    using V = typename ValT<VB>::type;
    constexpr int TILE = kSortThreads * IPT;
    constexpr int SB = (int)sizeof(KB) > VB ? (int)sizeof(KB) : VB;
    __shared__ uint32_t s_cnt[4][256];
    __shared__ uint32_t s_goff[256];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_tile;
    __shared__ bool s_last;
    __shared__ __attribute__((aligned(16))) unsigned char s_stage[TILE * SB];
    KB *s_keys = (KB *)s_stage;
    V *s_vals = (V *)s_stage;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t epoch = launch_epoch(ctrl);
    if (tid == 0) s_tile = next_tile(ctrl);
    for (int i = tid; i < 4 * 256; i += kSortThreads) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)TILE;
    const bool full = (uint64_t)base + TILE <= n;
    const uint32_t wb = base + w * 64 * IPT + lane;

    // wave-striped load: lane l of wave w holds tile positions w*64*IPT + j*64 + l
    KB k[IPT];
    V v[IPT];
    uint32_t d[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        uint32_t idx = wb + j * 64;
        bool ok = full || idx < n;
        k[j] = ok ? kin[idx] : KB(0);
        if constexpr (VB > 0) {
            if (ok) v[j] = vin[idx];
        }
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) d[j] = digit_of(k[j], xf, shift);

    // rank within the wave: peers = lanes whose digit matches (8 ballots);
    // rank = wave counter + peers below; the lowest peer bumps the counter.
    uint32_t r[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool ok = full || (wb + j * 64 < n);
        uint64_t peers = full ? ~0ull : ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint32_t nb = ((d[j] >> b) & 1u) - 1u;  // 0 if the bit is set, ~0 if clear
            uint64_t m = ballot(nb == 0u);
            uint64_t sel = m ^ (((uint64_t)nb << 32) | nb);
            peers &= sel;
        }
        uint32_t below = mbcnt(peers);
        uint32_t old = s_cnt[w][d[j]];
        r[j] = old + below;
        if (ok && below == 0) s_cnt[w][d[j]] = old + popc64(peers);
    }
    __syncthreads();

    // per digit (thread tid == digit): wave prefixes and the tile count
    const uint32_t c0 = s_cnt[0][tid], c1 = s_cnt[1][tid], c2 = s_cnt[2][tid], c3 = s_cnt[3][tid];
    const uint32_t total = c0 + c1 + c2 + c3;
    if (tile == 0) lb_store(&status[tid], lb_word(epoch, LB_P, total));
    else lb_store(&status[(size_t)tile * 256 + tid], lb_word(epoch, LB_A, total));
    uint32_t inc = wave_inclusive_scan(total, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t lstart = inc - total;
    for (uint32_t i = 0; i < w; ++i) lstart += s_wsum[i];
    s_cnt[0][tid] = lstart;
    s_cnt[1][tid] = lstart + c0;
    s_cnt[2][tid] = lstart + c0 + c1;
    s_cnt[3][tid] = lstart + c0 + c1 + c2;
    // decoupled look-back for this digit
    uint32_t excl = 0;
    if (tile > 0) {
        size_t kt = (size_t)tile - 1;
        uint32_t spins = 0;
        while (true) {
            uint64_t wd = lb_load(&status[kt * 256 + tid]);
            if (lb_epoch(wd) != epoch) {
                if (!spin_expired(ctrl, spins)) continue;
                break;
            }
            excl += lb_payload(wd);
            if (lb_flag(wd) == LB_P || kt == 0) break;
            --kt;
        }
        lb_store(&status[(size_t)tile * 256 + tid], lb_word(epoch, LB_P, excl + total));
    }
    s_goff[tid] = ctrl->offsets[pass][tid] + excl - lstart;
    __syncthreads();

    // reorder the tile by digit in LDS, then write out coalesced
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (full || wb + j * 64 < n) s_keys[s_cnt[w][d[j]] + r[j]] = k[j];
    }
    __syncthreads();
    const uint32_t tile_n = full ? (uint32_t)TILE : n - base;
    uint32_t od[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        uint32_t i = j * kSortThreads + tid;
        if (i < tile_n) {
            KB key = s_keys[i];
            od[j] = digit_of(key, xf, shift);
            uint32_t dst = s_goff[od[j]] + i;
            if (dst < n) kout[dst] = key;
            else __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if constexpr (VB > 0) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (full || wb + j * 64 < n) s_vals[s_cnt[w][d[j]] + r[j]] = v[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            uint32_t i = j * kSortThreads + tid;
            if (i < tile_n) {
                uint32_t dst = s_goff[od[j]] + i;
                if (dst < n) vout[dst] = s_vals[i];
            }
        }
    }
    if (arrive_last(ctrl, gridDim.x, &s_last) && tid == 0) finish_launch(ctrl, true);
}

// ------------------------------------------------------- insertion sort
// serial_insertion_sort(_by_key) (insertion_sort.hpp:42-64, :106-133): one
// work-item, shift while compare(value, data[pos-1]) with operator< (less)
// or operator> (greater) on the native key type.
template <typename T, int VB>
__global__ void k_insertion(T *keys, typename ValT<VB>::type *vals, uint32_t n, int descending)
{
    using V = typename ValT<VB>::type;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (uint32_t i = 1; i < n; ++i) {
        const T key = keys[i];
        V val{};
        if constexpr (VB > 0) val = vals[i];
        uint32_t pos = i;
        while (pos > 0 && (descending ? (key > keys[pos - 1]) : (key < keys[pos - 1]))) {
            keys[pos] = keys[pos - 1];
            if constexpr (VB > 0) vals[pos] = vals[pos - 1];
            pos--;
        }
        keys[pos] = key;
        if constexpr (VB > 0) vals[pos] = val;
    }
}

template <typename KB>
static KeyXf<KB> make_xf(int kind, int descending)
{
    KeyXf<KB> f;
    f.neg = 0;
    f.flt = 0;
    f.cx = 0;
    const KB sgn = (KB)((KB)1 << (sizeof(KB) * 8 - 1));
    if (!descending) {
        if (kind == 1) f.cx = sgn;
        if (kind == 2) f.flt = 1;
    } else {
        if (kind == 0) f.cx = (KB)~(KB)0;
        if (kind == 1) { f.neg = 1; f.cx = sgn; }
        if (kind == 2) { f.neg = 1; f.flt = 1; }
    }
    return f;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename KB, int VB, int IPT>
static int radix_sort(bcmi_queue q, void *keys_, void *vals_, uint32_t n, int kind, int descending)
{
    using V = typename ValT<VB>::type;
    constexpr int P = sizeof(KB);
    constexpr uint32_t TILE = kSortThreads * IPT;
    KB *keys = (KB *)keys_;
    V *vals = (V *)vals_;
    const KeyXf<KB> xf = make_xf<KB>(kind, descending);
    const uint32_t tiles = (n + TILE - 1) / TILE;
    void *scr;
    const size_t kbytes = align256((size_t)n * sizeof(KB));
    BCMI_TRY(scratch(q, kbytes + (VB ? (size_t)n * VB : 0) + 256, &scr));
    KB *ktmp = (KB *)scr;
    V *vtmp = (V *)((char *)scr + kbytes);
    uint64_t *status;
    BCMI_TRY(status_array(q, (size_t)tiles * 256, &status));

    uint32_t hblocks = (n + kHistThreads * 16 - 1) / (kHistThreads * 16);
    uint32_t hmax = (uint32_t)q->cu_count * 2;
    if (hblocks > hmax) hblocks = hmax;
    if (hblocks < 1) hblocks = 1;
    BCMI_LAUNCH(q, "radix_histogram",
                hipLaunchKernelGGL((k_histogram<KB>), dim3(hblocks), dim3(kHistThreads), 0, q->stream, keys, n, xf,
                                   q->ctrl));
    KB *ks = keys, *kd = ktmp;
    V *vs = vals, *vd = vtmp;
    for (int p = 0; p < P; ++p) {
        BCMI_LAUNCH(q, "radix_onesweep",
                    hipLaunchKernelGGL((k_onesweep<KB, VB, IPT>), dim3(tiles), dim3(kSortThreads), 0, q->stream, ks,
                                       kd, vs, vd, n, (uint32_t)(8 * p), (uint32_t)p, xf, q->ctrl, status));
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    BCMI_TRY(note_lookback_launch(q, P));
    if (ks != keys) {
        // odd number of passes (8-bit keys): the result is in the temporary
        BCMI_HIP(hipMemcpyAsync(keys, ks, (size_t)n * sizeof(KB), hipMemcpyDeviceToDevice, q->stream));
        if (VB) BCMI_HIP(hipMemcpyAsync(vals, vs, (size_t)n * VB, hipMemcpyDeviceToDevice, q->stream));
    }
    return 0;
}

template <typename KB, int VB>
static int radix_sort_ipt(bcmi_queue q, void *keys, void *vals, uint32_t n, int kind, int desc)
{
    constexpr int SB = (int)sizeof(KB) > VB ? (int)sizeof(KB) : VB;
    constexpr int IPT = SB >= 16 ? 8 : (SB >= 8 ? 12 : 16);
    return radix_sort<KB, VB, IPT>(q, keys, vals, n, kind, desc);
}

template <typename KB>
static int radix_sort_vb(bcmi_queue q, void *keys, void *vals, size_t vb, uint32_t n, int kind, int desc)
{
    if (!vals) vb = 0;
    switch (vb) {
    case 0: return radix_sort_ipt<KB, 0>(q, keys, nullptr, n, kind, desc);
    case 1: return radix_sort_ipt<KB, 1>(q, keys, vals, n, kind, desc);
    case 2: return radix_sort_ipt<KB, 2>(q, keys, vals, n, kind, desc);
    case 4: return radix_sort_ipt<KB, 4>(q, keys, vals, n, kind, desc);
    case 8: return radix_sort_ipt<KB, 8>(q, keys, vals, n, kind, desc);
    case 16: return radix_sort_ipt<KB, 16>(q, keys, vals, n, kind, desc);
    default: set_error("sort: value size %zu unsupported (1, 2, 4, 8 or 16 bytes)", vb); return BCMI_INVALID_VALUE;
    }
}

template <typename T, int VB>
static int insertion_launch(bcmi_queue q, void *keys, void *vals, uint32_t n, int desc)
{
    BCMI_LAUNCH(q, "insertion_sort",
                hipLaunchKernelGGL((k_insertion<T, VB>), dim3(1), dim3(64), 0, q->stream, (T *)keys,
                                   (typename ValT<VB>::type *)vals, n, desc));
    return 0;
}

template <typename T>
static int insertion_vb(bcmi_queue q, void *keys, void *vals, size_t vb, uint32_t n, int desc)
{
    if (!vals) vb = 0;
    switch (vb) {
    case 0: return insertion_launch<T, 0>(q, keys, nullptr, n, desc);
    case 1: return insertion_launch<T, 1>(q, keys, vals, n, desc);
    case 2: return insertion_launch<T, 2>(q, keys, vals, n, desc);
    case 4: return insertion_launch<T, 4>(q, keys, vals, n, desc);
    case 8: return insertion_launch<T, 8>(q, keys, vals, n, desc);
    case 16: return insertion_launch<T, 16>(q, keys, vals, n, desc);
    default: set_error("sort: value size %zu unsupported (1, 2, 4, 8 or 16 bytes)", vb); return BCMI_INVALID_VALUE;
    }
}

static int dtype_kind(int dt)
{
    switch (dt) {
    case BCMI_I8: case BCMI_I16: case BCMI_I32: case BCMI_I64: return 1;
    case BCMI_F32: case BCMI_F64: return 2;
    default: return 0;
    }
}

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_sort(bcmi_queue q, int api, int key_dtype, void *keys, void *values, size_t value_bytes, size_t n,
                         int descending)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(dtype_valid(key_dtype), BCMI_INVALID_VALUE, "sort: bad key dtype %d", key_dtype);
    BCMI_CHECK(api >= BCMI_SORT && api <= BCMI_RADIX_SORT, BCMI_INVALID_VALUE, "sort: bad api %d", api);
    // 32-bit tile/offset arithmetic (the reference's counts are uint too,
    // radix_sort.hpp:131-137); the margin keeps the last tile in range.
    BCMI_CHECK(n <= (1ull << 32) - (1ull << 20), BCMI_INVALID_BUFFER_SIZE, "sort: n=%zu exceeds 2^32-2^20", n);
    if (n < 2) return 0;
    BCMI_CHECK(keys, BCMI_INVALID_VALUE, "sort: null keys");
    if (api == BCMI_SORT || api == BCMI_STABLE_SORT) values = nullptr;
    const uint32_t n32 = (uint32_t)n;
    const int desc = descending ? 1 : 0;
    const bool insertion = (api == BCMI_SORT && n <= 32) ||
                           ((api == BCMI_SORT_BY_KEY || api == BCMI_STABLE_SORT_BY_KEY) && n < 32);
    if (insertion) {
        switch (key_dtype) {
        case BCMI_I8: return insertion_vb<int8_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_U8: return insertion_vb<uint8_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_I16: return insertion_vb<int16_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_U16: return insertion_vb<uint16_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_I32: return insertion_vb<int32_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_U32: return insertion_vb<uint32_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_I64: return insertion_vb<int64_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_U64: return insertion_vb<uint64_t>(q, keys, values, value_bytes, n32, desc);
        case BCMI_F32: return insertion_vb<float>(q, keys, values, value_bytes, n32, desc);
        default: return insertion_vb<double>(q, keys, values, value_bytes, n32, desc);
        }
    }
    const int kind = dtype_kind(key_dtype);
    switch (dtype_size(key_dtype)) {
    case 1: return radix_sort_vb<uint8_t>(q, keys, values, value_bytes, n32, kind, desc);
    case 2: return radix_sort_vb<uint16_t>(q, keys, values, value_bytes, n32, kind, desc);
    case 4: return radix_sort_vb<uint32_t>(q, keys, values, value_bytes, n32, kind, desc);
    default: return radix_sort_vb<uint64_t>(q, keys, values, value_bytes, n32, kind, desc);
    }
}
