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
// kernel per pass: block b takes tile b, ranks its keys with
// wave64 ballot matching against per-wave LDS digit counters, publishes its
// per-digit counts, gets its global digit offsets by decoupled look-back,
// reorders the tile in LDS so each digit run is contiguous and writes keys
// (then values) out coalesced.  The order equals the reference's: a stable
// sort by the transformed key radix(x) (radix_sort.hpp:98-129).
#include "onesweep.hpp"

#include <cstring>
#include <mutex>

namespace bcmi {

// ------------------------------------------------------------ histogram
// All passes' digit histograms in one read of the keys.  Each block counts
// its share in LDS and writes its partial histogram; k_digit_offsets then
// sums the partials per digit and scans them into exclusive digit offsets
// (ctrl->offsets).  No global atomics: device-scope atomics on shared
// addresses serialise across XCDs.
template <typename KB>
struct alignas(16) kchunk {
    static constexpr int N = 16 / sizeof(KB);
    KB v[N];
};

// Each group of L = kHistThreads / NW lanes counts into its own LDS
// sub-histogram: lanes adding to one bin in the same instruction serialise,
// which matters for skewed digits (the few top-digit values of floats in a
// narrow range); keys are read as 16-byte chunks, U per step, with the next
// step's U chunks in flight while a step is counted.
//
// BG = 0: sub-histogram s is 256 consecutive words per digit position, so
// bin d sits in bank d mod 64 for every sub and all 64 lanes of a wave
// contend for the banks.  BG = 1 (bank groups): the 64 / L subs of a wave
// interleave their bins so that sub sw uses only banks [sw * L, sw * L + L)
// (bin d at row d / L, column sw * L + d mod L of a 64-word row): the subs
// of a wave never contend, each sub's L lanes contend among L banks.
template <typename KB, typename XF, int U, int NW, int BG>
__global__ __launch_bounds__(kHistThreads) void k_histogram(const KB *__restrict__ keys, uint32_t n, uint32_t head,
                                                            XF xf, uint32_t *__restrict__ partial)
{
    constexpr int P = sizeof(KB);
    constexpr int L = kHistThreads / NW;  // lanes per sub
    static_assert(L <= kWave && kWave % L == 0, "subs within one wave");
    __shared__ uint32_t s_hist[NW * P * 256];
    for (int i = threadIdx.x; i < NW * P * 256; i += kHistThreads) s_hist[i] = 0;
    block_barrier();
    const uint32_t sub = threadIdx.x / L;
    // word of (position p, bin d) for this lane's sub
    auto slot = [&](int p, uint32_t d) -> uint32_t {
        if constexpr (BG) {
            constexpr uint32_t SPW = kWave / L;  // subs per wave
            const uint32_t w = sub / SPW, sw = sub % SPW;
            return ((w * P + (uint32_t)p) * SPW * 256) + (d / L) * kWave + sw * L + d % L;
        } else {
            return (sub * P + (uint32_t)p) * 256 + d;
        }
    };
    auto count = [&](KB x) {
        KB k = xf.key(x);
#pragma unroll
        for (int p = 0; p < P; ++p) atomicAdd(&s_hist[slot(p, (uint32_t)(k >> (8 * p)) & 0xFF)], 1u);
    };
    // unaligned head and the tail: block 0
    const uint64_t nchunks = (n - head) / kchunk<KB>::N;
    const uint64_t tail0 = head + nchunks * kchunk<KB>::N;
    if (blockIdx.x == 0) {
        for (uint32_t i = threadIdx.x; i < head; i += kHistThreads) count(keys[i]);
        for (uint64_t i = tail0 + threadIdx.x; i < n; i += kHistThreads) count(keys[i]);
    }
    const kchunk<KB> *body = (const kchunk<KB> *)(keys + head);
    const uint64_t stride = (uint64_t)gridDim.x * kHistThreads;
    uint64_t c = (uint64_t)blockIdx.x * kHistThreads + threadIdx.x;
    // U chunks per step, the next step's chunks loaded before this step's
    // are counted (2U in flight while the LDS works)
    if (c + (U - 1) * stride < nchunks) {
        kchunk<KB> cur[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cur[u] = body[c + u * stride];
        while (true) {
            const uint64_t nc = c + U * stride;
            const bool more = nc + (U - 1) * stride < nchunks;
            kchunk<KB> nxt[U];
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) nxt[u] = body[nc + u * stride];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int k = 0; k < kchunk<KB>::N; ++k) count(cur[u].v[k]);
            c = nc;
            if (!more) break;
#pragma unroll
            for (int u = 0; u < U; ++u) cur[u] = nxt[u];
        }
    }
    for (; c < nchunks; c += stride) {
        const kchunk<KB> a = body[c];
#pragma unroll
        for (int k = 0; k < kchunk<KB>::N; ++k) count(a.v[k]);
    }
    block_barrier();
    uint32_t *dst = partial + (size_t)blockIdx.x * P * 256;
    for (int i = threadIdx.x; i < P * 256; i += kHistThreads) {
        const uint32_t p = (uint32_t)i / 256, d = (uint32_t)i % 256;
        uint32_t t = 0;
#pragma unroll
        for (uint32_t s = 0; s < (uint32_t)NW; ++s) {
            if constexpr (BG) {
                constexpr uint32_t SPW = kWave / L;
                t += s_hist[((s / SPW * P + p) * SPW * 256) + (d / L) * kWave + (s % SPW) * L + d % L];
            } else {
                t += s_hist[(s * P + p) * 256 + d];
            }
        }
        dst[i] = t;
    }
}

// The histogram's shape: chunks in flight per thread, sub-histograms per
// block (LDS: subs * key bytes KB), blocks per CU, bank-grouped layout.
// Tuning cache keys "__bcmi_radix_hist" and, for sorts with values,
// "__bcmi_radix_hist_by_key"; params "unroll" (1, 2, 4, 8), "subs"
// (8 or 16; 8 for 8-byte keys), "blocks_per_cu" and "bank_groups" (0, 1);
// the defaults were the fastest measured on 2^27 uint32 keys.
struct HistShape {
    int unroll, subs, blocks_per_cu, bank_groups;
};

static HistShape hist_shape(size_t key_bytes, bool tuned, bool values)
{
    HistShape h{4, 16, 2, 0};
    if (tuned) {
        // sorts with values read "__bcmi_radix_hist_by_key" where it is set
        // (their benchmark's float keys have a skewed top digit)
        const char *k0 = "__bcmi_radix_hist", *k1 = values ? "__bcmi_radix_hist_by_key" : k0;
        h.unroll = (int)tune_get(k1, "unroll", tune_get(k0, "unroll", 4));
        h.subs = (int)tune_get(k1, "subs", tune_get(k0, "subs", 16));
        h.blocks_per_cu = (int)tune_get(k1, "blocks_per_cu", tune_get(k0, "blocks_per_cu", 2));
        h.bank_groups = tune_get(k1, "bank_groups", tune_get(k0, "bank_groups", 0)) ? 1 : 0;
    }
    if (h.unroll != 1 && h.unroll != 2 && h.unroll != 8) h.unroll = 4;
    if (h.subs != 8 || key_bytes >= 8) h.subs = key_bytes >= 8 ? 8 : 16;
    if (h.blocks_per_cu < 1 || h.blocks_per_cu > 8) h.blocks_per_cu = 2;
    return h;
}

template <typename KB, typename XF>
static int launch_histogram(bcmi_queue q, const HistShape &hs, uint32_t hblocks, const KB *keys, uint32_t n,
                            uint32_t head, XF xf, uint32_t *partial)
{
#define BCMI_HIST(U, NW, BG)                                                                                     \
    if (hs.unroll == U && hs.subs == NW && hs.bank_groups == BG) {                                              \
        BCMI_LAUNCH(q, "radix_histogram",                                                                       \
                    hipLaunchKernelGGL((k_histogram<KB, XF, U, NW, BG>), dim3(hblocks), dim3(kHistThreads), 0,  \
                                       q->stream, keys, n, head, xf, partial));                                 \
        return 0;                                                                                               \
    }
    if constexpr (sizeof(KB) < 8) {
        BCMI_HIST(1, 16, 0) BCMI_HIST(2, 16, 0) BCMI_HIST(4, 16, 0) BCMI_HIST(8, 16, 0)
        BCMI_HIST(1, 16, 1) BCMI_HIST(2, 16, 1) BCMI_HIST(4, 16, 1) BCMI_HIST(8, 16, 1)
    }
    BCMI_HIST(1, 8, 0) BCMI_HIST(2, 8, 0) BCMI_HIST(4, 8, 0) BCMI_HIST(8, 8, 0)
    BCMI_HIST(1, 8, 1) BCMI_HIST(2, 8, 1) BCMI_HIST(4, 8, 1) BCMI_HIST(8, 8, 1)
#undef BCMI_HIST
    set_error("sort: histogram shape unroll=%d subs=%d not built", hs.unroll, hs.subs);
    return BCMI_INVALID_VALUE;
}

// One block of 1024 threads per pass: thread (slice, d) sums digit d over a
// quarter of the partial histograms, the slices are combined in LDS and the
// 256 totals scanned into exclusive digit offsets.
__global__ __launch_bounds__(1024) void k_digit_offsets(const uint32_t *__restrict__ partial, uint32_t nparts,
                                                        uint32_t passes, Ctrl *ctrl)
{
    __shared__ uint32_t s_part[4][256];
    __shared__ uint32_t s_wsum[4];
    const uint32_t p = blockIdx.x, d = threadIdx.x & 255, slice = threadIdx.x >> 8;
    uint32_t c = 0;
#pragma unroll 8
    for (uint32_t b = slice; b < nparts; b += 4) c += partial[((size_t)b * passes + p) * 256 + d];
    s_part[slice][d] = c;
    block_barrier();
    const uint32_t lane = d & 63, w = d >> 6;
    c = s_part[0][d] + s_part[1][d] + s_part[2][d] + s_part[3][d];
    uint32_t inc = wave_inclusive_scan(c, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if (slice == 0 && lane == 63) s_wsum[w] = inc;
    block_barrier();
    if (slice == 0) {
        uint32_t pre = 0;
        for (uint32_t i = 0; i < w; ++i) pre += s_wsum[i];
        ctrl->offsets[p][d] = pre + inc - c;
    }
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

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace of one radix sort: the queue's arena and status array (base ==
// nullptr), or a fixed region the caller hands in (the allocation-free entry
// point's device scratch), carved into [status words | partial histograms |
// key temporaries | value temporaries].
struct SortWork {
    unsigned char *base;
    size_t bytes;
};

// Device scratch of bcmi_sort_noalloc: one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (in
// host code the name is a host-side stand-in); it uses the device address
// hipGetSymbolAddress returned for the current device, kept per device.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace bcmi
static __device__ __align__(256) unsigned char scratch[bcmi::kNoAllocScratchBytes];
namespace bcmi {
static unsigned char *g_noalloc_ptr[kNoAllocMaxDevices];
static std::mutex g_noalloc_mu;

// Called by bcmi_queue_create with `device` current: looks up this device's
// copy of the scratch array once.
int noalloc_scratch_setup(int device)
{
    BCMI_CHECK(device >= 0 && device < kNoAllocMaxDevices, BCMI_DEVICE_NOT_FOUND,
               "device %d beyond the %d devices the allocation-free sort supports", device, kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return 0;
    void *p = nullptr;
    BCMI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    BCMI_CHECK(p != nullptr, BCMI_RUNTIME_ERROR, "no device address for the sort scratch on device %d", device);
    g_noalloc_ptr[device] = (unsigned char *)p;
    return 0;
}

static unsigned char *noalloc_scratch(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

__global__ void k_clear_words(uint64_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}

// Bytes of queue scratch radix_sort_xf takes for n keys of key_bytes with
// value_bytes values (at most cu_count * 8 histogram blocks).  Callers that
// keep their own temporaries in the arena beside a sort place them after it.
size_t radix_scratch_bytes(size_t n, size_t key_bytes, size_t value_bytes, int cu_count)
{
    return align256(n * key_bytes) + align256(n * value_bytes) + align256((size_t)cu_count * 8 * key_bytes * 256 * 4);
}

template <typename KB, int VB, int TH, int IPT, int SPLIT, int LB, int WPE, int CBG, typename XF>
static int radix_sort_xf(bcmi_queue q, void *keys_, void *vals_, uint32_t n, XF xf, const SortWork *work,
                         unsigned xrun)
{
    using V = typename ValT<VB>::type;
    constexpr int P = sizeof(KB);
    constexpr uint32_t TILE = TH * IPT;
    KB *keys = (KB *)keys_;
    V *vals = (V *)vals_;
    const uint32_t tiles = (n + TILE - 1) / TILE;
    // XCD-grouped tile order only where the device holds at least twice a
    // group of 8 * xrun workgroups at once (LDS and thread limits of a CU)
    {
        const size_t per_cu_lds = (size_t)kLdsBytesPerCU / sizeof(QSweepLds<TH, IPT, KB, VB, SPLIT>);
        const size_t per_cu = per_cu_lds < (size_t)(2048 / TH) ? per_cu_lds : (size_t)(2048 / TH);
        if (xrun > 16 || (size_t)q->cu_count * per_cu < 16 * (size_t)xrun) xrun = 0;
    }
    const HistShape hs = hist_shape(sizeof(KB), work == nullptr, VB > 0);
    uint32_t hblocks = (n + kHistThreads * 32 - 1) / (kHistThreads * 32);
    uint32_t hmax = (uint32_t)(q->cu_count * hs.blocks_per_cu);
    if (hblocks > hmax) hblocks = hmax;
    if (hblocks < 1) hblocks = 1;
    uint32_t head = 0;
    if ((uintptr_t)keys & 15) head = (uint32_t)((16 - ((uintptr_t)keys & 15)) / sizeof(KB));
    if (((uintptr_t)keys & 15) % sizeof(KB)) head = n;  // not element-aligned to a chunk
    if (head > n) head = n;
    const size_t kbytes = align256((size_t)n * sizeof(KB));
    const size_t vbytes = align256(VB ? (size_t)n * VB : 0);
    const size_t pbytes = align256((size_t)hblocks * P * 256 * 4);
    const size_t swords = (size_t)tiles * 256;
    unsigned char *scr;
    uint64_t *status;
    if (work) {
        const size_t need = align256(swords * 8) + pbytes + kbytes + vbytes;
        BCMI_CHECK(need <= work->bytes, BCMI_INVALID_BUFFER_SIZE,
                   "sort_noalloc: %u keys of %zu bytes with %d-byte values need %zu bytes of scratch; "
                   "the allocation-free sort has %zu", n, sizeof(KB), VB, need, work->bytes);
        status = (uint64_t *)work->base;
        scr = work->base + align256(swords * 8);
        // the words may carry another queue's epochs: clear them (a fill
        // kernel, so the call stays capturable)
        BCMI_LAUNCH(q, "radix_clear_status",
                    hipLaunchKernelGGL(k_clear_words, dim3((unsigned)((swords + 255) / 256 < 4096 ? (swords + 255) / 256 : 4096)),
                                       dim3(256), 0, q->stream, status, swords));
    } else {
        void *a;
        BCMI_TRY(scratch(q, pbytes + kbytes + vbytes, &a));
        scr = (unsigned char *)a;
        BCMI_TRY(status_array(q, swords, &status));
    }
    uint32_t *partial = (uint32_t *)scr;
    KB *ktmp = (KB *)(scr + pbytes);
    V *vtmp = (V *)(scr + pbytes + kbytes);

    BCMI_TRY(launch_histogram(q, hs, hblocks, keys, n, head, xf, partial));
    BCMI_LAUNCH(q, "radix_digit_offsets",
                hipLaunchKernelGGL(k_digit_offsets, dim3(P), dim3(1024), 0, q->stream, partial, hblocks,
                                   (uint32_t)P, q->ctrl));
    BCMI_TRY(begin_lookback(q, P));
    KB *ks = keys, *kd = ktmp;
    V *vs = vals, *vd = vtmp;
    for (int p = 0; p < P; ++p) {
        BCMI_LAUNCH(q, "radix_onesweep",
                    hipLaunchKernelGGL((k_qsweep<TH, IPT, KB, VB, SPLIT, LB, WPE, XF, CBG>), dim3(tiles), dim3(TH), 0, q->stream, ks, kd,
                                       vs, vd, n, (uint32_t)(8 * p), (uint32_t)p, xf, q->ctrl, status,
                                       (uint32_t)(P - p), (uint32_t)xrun));
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    if (ks != keys) {
        // odd number of passes (8-bit keys): the result is in the temporary
        BCMI_HIP(hipMemcpyAsync(keys, ks, (size_t)n * sizeof(KB), hipMemcpyDeviceToDevice, q->stream));
        if (VB) BCMI_HIP(hipMemcpyAsync(vals, vs, (size_t)n * VB, hipMemcpyDeviceToDevice, q->stream));
    }
    return 0;
}

template <typename KB, int VB, int TH, int IPT, int SPLIT = 1, int LB = 4, int WPE = 1, int CBG = 0>
static int radix_sort(bcmi_queue q, void *keys, void *vals, uint32_t n, int kind, int descending,
                      const SortWork *work, unsigned xrun)
{
    return with_key_xf<KB>(kind, descending, [&](auto xf) {
        return radix_sort_xf<KB, VB, TH, IPT, SPLIT, LB, WPE, CBG>(q, keys, vals, n, xf, work, xrun);
    });
}

// Tile shape: threads per workgroup, keys per thread and staging rounds.
// For 4-byte keys with no or 4-byte values it comes from the tuning cache
// (key "__bcmi_radix_sort", or "__bcmi_radix_sort_by_key" with values;
// params "threads", "ipt", "split", "lookback" (predecessor words per
// look-back round trip) and "waves" (a register bound, waves per SIMD; 0 =
// none), "count_bg" (1: bank-grouped count words, built for the best
// shapes), chosen on 2^27 keys; the lookback / waves combinations built are
// listed below; smaller sorts take
// smaller tiles to fill the chip (param "fit_size" = 0 keeps the tuned shape
// at every size).  Other key and value widths take a fixed shape whose
// staging buffer stays within 32-64 KB (two workgroups per CU).
template <typename KB, int VB>
static int radix_sort_shape(bcmi_queue q, void *keys, void *vals, uint32_t n, int kind, int desc,
                            const SortWork *work)
{
    constexpr int SB = (int)sizeof(KB) > VB ? (int)sizeof(KB) : VB;
    // the allocation-free sort (work != nullptr) reads no tuning cache
    const char *key = VB ? "__bcmi_radix_sort_by_key" : "__bcmi_radix_sort";
    const unsigned xrun = work ? 0u : tune_get(key, "xcd_run", 0);
    if constexpr (sizeof(KB) == 4 && (VB == 0 || VB == 4)) {
        unsigned th = work ? 256 : tune_get(key, "threads", 512);
        unsigned ipt = work ? 32 : tune_get(key, "ipt", 32);
        unsigned split = work ? 1 : tune_get(key, "split", 1);
        const unsigned lb = work ? 4 : tune_get(key, "lookback", 4);
        const unsigned wpe = work ? 0 : tune_get(key, "waves", 0);
        const unsigned cbg = work ? 0 : tune_get(key, "count_bg", 0);
        bool tuned_shape = true;
        if (work || tune_get(key, "fit_size", 1) != 0) {
            if (n <= (1u << 20)) {
                th = 256;
                ipt = 16;
                split = 1;
                tuned_shape = false;
            } else if (n <= (1u << 23)) {
                th = 256;
                ipt = 32;
                split = 1;
                tuned_shape = false;
            }
        }
        // register bounds (param "waves": waves per SIMD) that raise the
        // workgroups per CU to what the LDS allows without spills, and
        // look-back windows other than 4 words per round trip, for the
        // tuned shapes; any other combination takes the plain shape
#define BCMI_SHAPE_LB(T, I, S, W, E)                                                                        \
        if constexpr (sizeof(QSweepLds<T, I, KB, VB, S>) <= (size_t)kLdsBytesPerCU)                         \
            if (th == T && ipt == I && split == S && lb == W && wpe == E && cbg == 0)                      \
                return radix_sort<KB, VB, T, I, S, W, E>(q, keys, vals, n, kind, desc, work, xrun)
        // bank-grouped count words (param "count_bg" = 1) for the shapes
        // that tune best
#define BCMI_SHAPE_CB(T, I, S, W, E)                                                                        \
        if constexpr (sizeof(QSweepLds<T, I, KB, VB, S>) <= (size_t)kLdsBytesPerCU)                         \
            if (th == T && ipt == I && split == S && lb == W && wpe == E && cbg == 1)                      \
                return radix_sort<KB, VB, T, I, S, W, E, 1>(q, keys, vals, n, kind, desc, work, xrun)
        if (tuned_shape) {
            BCMI_SHAPE_CB(512, 64, 2, 2, 4);
            BCMI_SHAPE_CB(256, 64, 2, 2, 4);
            BCMI_SHAPE_CB(512, 32, 2, 2, 0);
            BCMI_SHAPE_CB(512, 32, 2, 2, 7);
            BCMI_SHAPE_LB(256, 64, 2, 2, 0);
            BCMI_SHAPE_LB(512, 32, 2, 2, 0);
            BCMI_SHAPE_LB(256, 64, 2, 2, 4);
            BCMI_SHAPE_LB(512, 24, 2, 2, 8);
            BCMI_SHAPE_LB(512, 32, 2, 2, 7);
            BCMI_SHAPE_LB(512, 16, 1, 4, 8);
            BCMI_SHAPE_LB(512, 64, 2, 2, 4);
            BCMI_SHAPE_LB(512, 64, 4, 2, 4);
            BCMI_SHAPE_LB(1024, 32, 2, 2, 0);
            BCMI_SHAPE_LB(1024, 32, 4, 2, 0);
        }
#undef BCMI_SHAPE_LB
#undef BCMI_SHAPE_CB
#define BCMI_SHAPE(T, I, S) \
        if (th == T && ipt == I && split == S) return radix_sort<KB, VB, T, I, S>(q, keys, vals, n, kind, desc, work, xrun)
        BCMI_SHAPE(256, 16, 1);
        BCMI_SHAPE(256, 32, 1);
        BCMI_SHAPE(256, 64, 1);
        BCMI_SHAPE(512, 16, 1);
        BCMI_SHAPE(512, 24, 1);
        BCMI_SHAPE(512, 32, 1);
        BCMI_SHAPE(1024, 16, 1);
        BCMI_SHAPE(256, 64, 2);
        BCMI_SHAPE(512, 24, 2);
        BCMI_SHAPE(512, 32, 2);
        BCMI_SHAPE(512, 24, 3);
#undef BCMI_SHAPE
        // any other value falls through to the generic shape below
    }
    if constexpr (SB >= 8) return radix_sort<KB, VB, 256, 16>(q, keys, vals, n, kind, desc, work, xrun);
    else return radix_sort<KB, VB, 256, 32>(q, keys, vals, n, kind, desc, work, xrun);
}

template <typename KB>
static int radix_sort_vb(bcmi_queue q, void *keys, void *vals, size_t vb, uint32_t n, int kind, int desc,
                         const SortWork *work)
{
    if (!vals) vb = 0;
    switch (vb) {
    case 0: return radix_sort_shape<KB, 0>(q, keys, nullptr, n, kind, desc, work);
    case 1: return radix_sort_shape<KB, 1>(q, keys, vals, n, kind, desc, work);
    case 2: return radix_sort_shape<KB, 2>(q, keys, vals, n, kind, desc, work);
    case 4: return radix_sort_shape<KB, 4>(q, keys, vals, n, kind, desc, work);
    case 8: return radix_sort_shape<KB, 8>(q, keys, vals, n, kind, desc, work);
    case 16: return radix_sort_shape<KB, 16>(q, keys, vals, n, kind, desc, work);
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

static int sort_entry(bcmi_queue q, int api, int key_dtype, void *keys, void *values, size_t value_bytes, size_t n,
                      int descending, const SortWork *work)
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
    case 1: return radix_sort_vb<uint8_t>(q, keys, values, value_bytes, n32, kind, desc, work);
    case 2: return radix_sort_vb<uint16_t>(q, keys, values, value_bytes, n32, kind, desc, work);
    case 4: return radix_sort_vb<uint32_t>(q, keys, values, value_bytes, n32, kind, desc, work);
    default: return radix_sort_vb<uint64_t>(q, keys, values, value_bytes, n32, kind, desc, work);
    }
}

extern "C" int bcmi_sort(bcmi_queue q, int api, int key_dtype, void *keys, void *values, size_t value_bytes, size_t n,
                         int descending)
{
    return sort_entry(q, api, key_dtype, keys, values, value_bytes, n, descending, nullptr);
}

// The same sort with every temporary in this device's copy of `scratch`:
// no allocation on any path.  Calls that overlap on one GPU (two queues at
// once) share the array and race; the array keeps the last call's contents.
extern "C" int bcmi_sort_noalloc(bcmi_queue q, int api, int key_dtype, void *keys, void *values, size_t value_bytes,
                                 size_t n, int descending)
{
    BCMI_TRY(use_queue(q));
    unsigned char *p = noalloc_scratch(q->device);
    BCMI_CHECK(p != nullptr, BCMI_INVALID_COMMAND_QUEUE, "sort_noalloc: no scratch set up for device %d", q->device);
    const SortWork w{p, kNoAllocScratchBytes};
    // the queue's kernel profiler creates its events on demand: off here
    const bool prof = q->profile;
    q->profile = false;
    const int r = sort_entry(q, api, key_dtype, keys, values, value_bytes, n, descending, &w);
    q->profile = prof;
    return r;
}
