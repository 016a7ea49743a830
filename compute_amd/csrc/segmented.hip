// segmented.hip -- batched segmented sort.
//
// Semantics: every segment [offsets[s], offsets[s+1]) ends up exactly as one
// stable_sort_by_key call on it would leave it in the reference's GPU
// column (stable_sort_by_key.hpp:29-77): below 32 elements the one-work-item
// insertion sort (insertion_sort.hpp:106-133, which only differs from radix
// order for float keys: it keeps -0.0/+0.0 in input order), otherwise the
// stable radix order of the radix() key transform (radix_sort.hpp:98-129).
//
// The reference has no batched primitive: the per-segment alternative is one
// sort call per segment (several launches each).  Here a plan built once
// from the host offsets routes the segments to three paths, each one launch
// (or one launch per 8-bit pass) for all segments together:
//   * small (<= 4096 elements): one workgroup per segment sorts it entirely
//     in LDS -- all digit passes, LDS-mask ranking, LDS scatter/gather --
//     so the segment's keys are read and written once;
//   * float keys shorter than 32: the reference's insertion sort, one
//     thread per segment;
//   * large: the onesweep pass kernel over a table of tiles of all large
//     segments, with per-segment digit histograms and look-back chains that
//     restart at every segment's first tile.
// A plan owns its tables and temporaries, so running it allocates nothing
// and can be captured in a hipGraph and replayed.
#include "onesweep.hpp"
#include "radix.hpp"
#include "segplan.hpp"

#include <cstring>
#include <vector>

namespace bcmi {

// ------------------------------------------------------------ small path
constexpr uint32_t kBucketMax = 64;  // largest bucket the one-pass sort finishes by insertion

// Sorts each span of segs entirely in LDS.  The keys (values) are read from
// ksrc (vsrc) and the sorted span written to keys (vals) at the same offset;
// the small-segment route sorts in place (ksrc == keys).
//
// First range-adaptive passes: with [mn, mx] the span's range of transformed
// keys, rel(k) = (key(k) - mn) >> shift, shift chosen so rel fits 8 bits
// (spans up to 512 keys) or 16 bits, is monotone in the key.  One or two
// stable 8-bit LDS passes order the keys by rel; then each run of equal rel
// (a few keys for spread-out keys: 4096 keys over 2^16 values) is finished
// by a stable insertion sort on the whole key, owned by the thread holding
// the run's first slot.  If a run exceeds kBucketMax keys (clustered keys)
// the span takes all 8-bit LSD passes in LDS instead.
template <typename KB, int VB, int IPT, typename XF>
__global__ __launch_bounds__(kSegSmallThreads) void k_seg_small(const KB *ksrc, const typename ValT<VB>::type *vsrc,
                                                                KB *keys, typename ValT<VB>::type *vals,
                                                                const SegSpan *__restrict__ segs, XF xf)
{
    using V = typename ValT<VB>::type;
    constexpr int P = sizeof(KB);
    constexpr int TILE = kSegSmallThreads * IPT;
    __shared__ KB s_k[TILE];
    __shared__ V s_v[VB ? TILE : 1];
    __shared__ uint32_t s_cnt[4][256];
    __shared__ uint64_t s_peer[4][256];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_uniform;
    __shared__ KB s_mn[4], s_mx[4];

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const SegSpan sg = segs[blockIdx.x];
    const uint32_t len = sg.len;
    if (len == 0) return;  // an unused group slot
    const KB *ks = ksrc + sg.base;
    const V *vs = VB ? vsrc + sg.base : nullptr;
    KB *kp = keys + sg.base;
    V *vp = VB ? vals + sg.base : nullptr;
    auto pos = [&](int j) { return w * 64 * IPT + j * 64 + lane; };
    auto valid = [&](int j) { return pos(j) < len; };
    const int32_t lim = (int32_t)len - (int32_t)(w * 64 * IPT);

    KB k[IPT];
    V v[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        k[j] = valid(j) ? ks[pos(j)] : KB(0);
        if constexpr (VB > 0) {
            if (valid(j)) v[j] = vs[pos(j)];
        }
    }
    for (int i = tid; i < 4 * 256; i += kSegSmallThreads) {
        (&s_peer[0][0])[i] = 0;
        (&s_cnt[0][0])[i] = 0;
    }

    // one stable LDS counting pass by dig() over the register-held keys:
    // in-wave ranks, the block's digit offsets, scatter to LDS, read back in
    // position order (a pass whose digit is the same for every key is skipped)
    auto lds_pass = [&](auto dig) {
        Ranks<IPT> r;
        for (int i = lane; i < 256; i += 64) s_cnt[w][i] = 0;
        block_barrier();
        // every thread read the previous pass's flag before the barrier above
        if (tid == 0) s_uniform = 0;
        rank_in_wave_by<IPT, false, kRankLdsMask>(k, r, dig, s_cnt[w], s_peer[w], lim);
        block_barrier();
        const uint32_t c0 = s_cnt[0][tid], c1 = s_cnt[1][tid], c2 = s_cnt[2][tid], c3 = s_cnt[3][tid];
        const uint32_t total = c0 + c1 + c2 + c3;
        if (total == len) s_uniform = 1;
        const uint32_t lstart = block_exclusive_sum256(total, s_wsum);
        s_cnt[0][tid] = lstart;
        s_cnt[1][tid] = lstart + c0;
        s_cnt[2][tid] = lstart + c0 + c1;
        s_cnt[3][tid] = lstart + c0 + c1 + c2;
        block_barrier();
        if (s_uniform) return;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (valid(j)) {
                const uint32_t dst = s_cnt[w][dig(k[j])] + r[j];
                s_k[dst] = k[j];
                if constexpr (VB > 0) s_v[dst] = v[j];
            }
        }
        block_barrier();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (valid(j)) {
                k[j] = s_k[pos(j)];
                if constexpr (VB > 0) v[j] = s_v[pos(j)];
            }
        }
    };

    // ---- range-adaptive passes
    {
        KB mn = (KB)~KB(0), mx = KB(0);
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (valid(j)) {
                const KB t = xf.key(k[j]);
                mn = t < mn ? t : mn;
                mx = t > mx ? t : mx;
            }
        }
        {  // wave reductions move 32-bit words: widen narrow keys
            using WK = std::conditional_t<(sizeof(KB) < 4), uint32_t, KB>;
            mn = (KB)wave_reduce((WK)mn, (WK)~WK(0), [](WK a, WK b) { return a < b ? a : b; });
            mx = (KB)wave_reduce((WK)mx, WK(0), [](WK a, WK b) { return a > b ? a : b; });
        }
        if (lane == 0) {
            s_mn[w] = mn;
            s_mx[w] = mx;
        }
        block_barrier();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            mn = s_mn[i] < mn ? s_mn[i] : mn;
            mx = s_mx[i] > mx ? s_mx[i] : mx;
        }
        if (mn == mx) {  // one key value: the input order is the sorted order
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if (valid(j)) {
                    kp[pos(j)] = k[j];
                    if constexpr (VB > 0) vp[pos(j)] = v[j];
                }
            }
            return;
        }
        // the leading 8 (spans up to 512 keys) or 16 bits of key - mn, as
        // one or two 8-bit passes, least significant first
        const uint32_t bits = len > 512 ? 16u : 8u;
        const uint32_t bitlen = 64u - (uint32_t)__builtin_clzll((unsigned long long)(KB)(mx - mn));
        const uint32_t shift = bitlen > bits ? bitlen - bits : 0u;
        auto rel = [&](KB x) { return (KB)((KB)(xf.key(x) - mn) >> shift); };  // < 2^bits
        if (bits == 16) lds_pass([&](KB x) { return (uint32_t)rel(x) & 0xFFu; });
        lds_pass([&](KB x) { return (uint32_t)(rel(x) >> (bits - 8)); });
        // keys are now ordered by rel(); finish each run of equal rel() with a
        // stable insertion sort on the whole key (owner: the run's first slot)
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (valid(j)) {
                s_k[pos(j)] = k[j];
                if constexpr (VB > 0) s_v[pos(j)] = v[j];
            }
        }
        block_barrier();
        uint32_t rlen[IPT];
        bool big = false;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            rlen[j] = 0;
            const uint32_t i = pos(j);
            if (!valid(j)) continue;
            const KB here = rel(s_k[i]);
            if (i > 0 && rel(s_k[i - 1]) == here) continue;  // not a run start
            uint32_t e = i + 1;
            while (e < len && e - i <= kBucketMax && rel(s_k[e]) == here) ++e;
            rlen[j] = e - i;
            big = big || rlen[j] > kBucketMax;
        }
        if (!block_barrier_or(big)) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t b0 = pos(j);
                for (uint32_t i = 1; i < rlen[j]; ++i) {
                    const KB key = s_k[b0 + i];
                    const KB tk = xf.key(key);
                    V val{};
                    if constexpr (VB > 0) val = s_v[b0 + i];
                    uint32_t q = i;
                    while (q > 0 && xf.key(s_k[b0 + q - 1]) > tk) {
                        s_k[b0 + q] = s_k[b0 + q - 1];
                        if constexpr (VB > 0) s_v[b0 + q] = s_v[b0 + q - 1];
                        --q;
                    }
                    s_k[b0 + q] = key;
                    if constexpr (VB > 0) s_v[b0 + q] = val;
                }
            }
            block_barrier();
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if (valid(j)) {
                    kp[pos(j)] = s_k[pos(j)];
                    if constexpr (VB > 0) vp[pos(j)] = s_v[pos(j)];
                }
            }
            return;
        }
        // a run too long to finish by insertion (clustered keys): all LSD
        // passes follow.  The order so far is a stable sort by a monotone
        // function of the key, so equal keys are still in input order.
    }

    for (int p = 0; p < P; ++p) lds_pass([&](KB x) { return xf.digit(x, 8 * p); });
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (valid(j)) {
            kp[pos(j)] = k[j];
            if constexpr (VB > 0) vp[pos(j)] = v[j];
        }
    }
}

// ------------------------------------------------- tiny float segments
// serial_insertion_sort_by_key (insertion_sort.hpp:106-133) on the native
// type, one thread per segment.
template <typename T, int VB>
__global__ void k_seg_insertion(T *keys, typename ValT<VB>::type *vals, const SegSpan *segs, uint32_t nsegs,
                                int descending)
{
    using V = typename ValT<VB>::type;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsegs) return;
    T *kp = keys + segs[s].base;
    V *vp = VB ? vals + segs[s].base : nullptr;
    const uint32_t n = segs[s].len;
    for (uint32_t i = 1; i < n; ++i) {
        const T key = kp[i];
        V val{};
        if constexpr (VB > 0) val = vp[i];
        uint32_t q = i;
        while (q > 0 && (descending ? (key > kp[q - 1]) : (key < kp[q - 1]))) {
            kp[q] = kp[q - 1];
            if constexpr (VB > 0) vp[q] = vp[q - 1];
            q--;
        }
        kp[q] = key;
        if constexpr (VB > 0) vp[q] = val;
    }
}

// ------------------------------------------------------------ large path
// Route of a large segment, decided on the device from its histogram:
//   kRouteLsd: all P onesweep passes over the segment (any distribution);
//   kRouteMsd: one onesweep pass on the most significant digit into the
//     temporaries, which partitions the segment stably into its 256 digit
//     buckets; consecutive buckets are packed greedily into groups of at
//     most `cap` keys, and each group is sorted in LDS by the whole key and
//     written back (a group is a contiguous range of the sorted segment, so
//     sorting it by the whole key is exact, and stable).  Taken when every
//     bucket fits one group and the groups fit the segment's slots.
enum : uint32_t { kRouteLsd = 0, kRouteMsd = 1 };

// One block per histogram chunk (<= kSegChunk keys of one segment): the
// histograms of digits [p0, p1).  With route != nullptr only chunks of
// segments on the LSD route are counted (the others need just the top digit).
template <typename KB, typename XF>
__global__ __launch_bounds__(kHistThreads) void k_seg_histogram(const KB *__restrict__ keys,
                                                                const SegSpan *__restrict__ chunks, XF xf,
                                                                uint32_t *__restrict__ partial, uint32_t p0,
                                                                uint32_t p1, const uint32_t *__restrict__ route)
{
    constexpr int P = sizeof(KB);
    constexpr int NW = kHistThreads / 64;
    __shared__ uint32_t s_hist[NW][P][256];
    const SegSpan c = chunks[blockIdx.x];
    if (route && route[c.idx] != kRouteLsd) return;
    for (int i = threadIdx.x; i < NW * P * 256; i += kHistThreads) (&s_hist[0][0][0])[i] = 0;
    block_barrier();
    uint32_t(*h)[256] = s_hist[threadIdx.x >> 6];
    const KB *kc = keys + c.base;
    constexpr int U = 8;  // loads in flight per thread
    for (uint32_t i0 = threadIdx.x; i0 < c.len; i0 += kHistThreads * U) {
        KB kk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * kHistThreads;
            kk[u] = i < c.len ? kc[i] : KB(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i0 + u * kHistThreads < c.len) {
                const KB k = xf.key(kk[u]);
                for (uint32_t p = p0; p < p1; ++p) atomicAdd(&h[p][(k >> (8 * p)) & 0xFF], 1u);
            }
        }
    }
    block_barrier();
    uint32_t *dst = partial + (size_t)blockIdx.x * P * 256;
    for (uint32_t i = p0 * 256 + threadIdx.x; i < p1 * 256; i += kHistThreads) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += (&s_hist[w][0][0])[i];
        dst[i] = t;
    }
}

// grid (large segments, passes p0 + blockIdx.y): digit d of segment s summed
// over its chunks [chunk_begin[s], chunk_begin[s+1]) and scanned into
// exclusive offsets; with route != nullptr, LSD-route segments only.
__global__ __launch_bounds__(256) void k_seg_digit_offsets(const uint32_t *__restrict__ partial,
                                                           const uint32_t *__restrict__ chunk_begin, uint32_t passes,
                                                           uint32_t p0, uint32_t *__restrict__ seg_off,
                                                           const uint32_t *__restrict__ route)
{
    __shared__ uint32_t s_wsum[4];
    const uint32_t s = blockIdx.x, p = p0 + blockIdx.y, d = threadIdx.x;
    if (route && route[s] != kRouteLsd) return;
    uint32_t c = 0;
    for (uint32_t b = chunk_begin[s]; b < chunk_begin[s + 1]; ++b) c += partial[((size_t)b * passes + p) * 256 + d];
    seg_off[((size_t)s * passes + p) * 256 + d] = block_exclusive_sum256(c, s_wsum);
}

template <typename KB, int VB, int IPT, typename XF>
__global__ __launch_bounds__(kSortThreads) __attribute__((amdgpu_waves_per_eu(onesweep_waves_per_eu<VB, IPT>())))
void k_onesweep_seg(const KB *__restrict__ kin, KB *__restrict__ kout,
                                                               const typename ValT<VB>::type *__restrict__ vin,
                                                               typename ValT<VB>::type *__restrict__ vout,
                                                               const SegTile *__restrict__ tiles, uint32_t shift,
                                                               uint32_t pass, uint32_t passes,
                                                               const uint32_t *__restrict__ seg_off, XF xf,
                                                               Ctrl *ctrl, uint64_t *__restrict__ status,
                                                               uint32_t back, uint32_t want)
{
    // the tile carries its segment's route (one load before a tile of the
    // other route exits)
    const SegTile t = tiles[blockIdx.x];
    if (t.route != want) return;  // the segment takes the other route
    const size_t b = t.base;
    onesweep_tile<KB, VB, IPT, kRankLdsMask, XF>(kin + b, kout + b, VB ? vin + b : vin, VB ? vout + b : vout, t.len,
                                             t.tile, shift, seg_off + ((size_t)t.seg * passes + pass) * 256, xf,
                                             ctrl, status + (size_t)(blockIdx.x - t.tile) * 256, back);
}

// The same pass on the quarter-wave digit-pass body (onesweep.hpp): 16-lane
// subs rank by OR-ed peer masks in bank-grouped count words, the tile is
// staged while its look-back completes.  Tile, look-back chain and offsets
// are the segment's, as above.
template <typename KB, int VB, int IPT, typename XF>
__global__ __launch_bounds__(kSortThreads) void k_qsweep_seg(const KB *__restrict__ kin, KB *__restrict__ kout,
                                                             const typename ValT<VB>::type *__restrict__ vin,
                                                             typename ValT<VB>::type *__restrict__ vout,
                                                             const SegTile *__restrict__ tiles, uint32_t shift,
                                                             uint32_t pass, uint32_t passes,
                                                             const uint32_t *__restrict__ seg_off, XF xf,
                                                             Ctrl *ctrl, uint64_t *__restrict__ status,
                                                             uint32_t back, uint32_t want)
{
    __shared__ QSweepLds<kSortThreads, IPT, KB, VB, 1> lds;
    const SegTile t = tiles[blockIdx.x];
    if (t.route != want) return;  // the segment takes the other route
    const uint32_t epoch = launch_epoch_plain(ctrl, back);
    const size_t b = t.base;
    const uint32_t *goff = seg_off + ((size_t)t.seg * passes + pass) * 256;
    uint64_t *st = status + (size_t)(blockIdx.x - t.tile) * 256;
    if ((uint64_t)(t.tile + 1) * (kSortThreads * IPT) <= t.len)
        qsweep_body<kSortThreads, IPT, KB, VB, 1, 2, XF, true, 1>(lds, kin + b, kout + b, VB ? vin + b : vin,
                                                                VB ? vout + b : vout, t.len, t.tile, shift, goff, xf,
                                                                ctrl, st, epoch);
    else
        qsweep_body<kSortThreads, IPT, KB, VB, 1, 2, XF, false, 1>(lds, kin + b, kout + b, VB ? vin + b : vin,
                                                                 VB ? vout + b : vout, t.len, t.tile, shift, goff,
                                                                 xf, ctrl, st, epoch);
}

// One block (256 threads, thread = digit) per large segment.  The top-digit
// buckets are packed greedily into groups of at most `cap` elements: a
// group that starts at bucket b ends before next(b), the first bucket whose
// end lies more than cap past b's start (a binary search over the bucket
// starts), and the group starts are the chain 0, next(0), next(next(0)), ...
// That chain is marked in 8 rounds of pointer doubling (it has at most 256
// nodes), so no thread walks the buckets one by one (a serial walk per
// segment is bound by the CU's one scalar unit).  A group's slot is its
// start's rank among the marks.  All threads then clear the segment's unused
// group slots and tag its tiles with the route.
__global__ __launch_bounds__(256) void k_seg_route(const SegSpan *__restrict__ large,
                                                   const uint32_t *__restrict__ seg_off, uint32_t passes,
                                                   const uint32_t *__restrict__ slot_begin, SegSpan *groups,
                                                   uint32_t *route, SegTile *tiles, uint32_t tile_len, uint32_t cap,
                                                   uint32_t enable)
{
    __shared__ uint32_t s_start[257];
    __shared__ uint32_t s_jump[2][257];
    __shared__ uint32_t s_mark[257];
    __shared__ uint32_t s_wave[4];
    const uint32_t s = blockIdx.x, d = threadIdx.x;
    const SegSpan sg = large[s];
    const uint32_t *off = seg_off + ((size_t)s * passes + passes - 1) * 256;
    s_start[d] = off[d];
    if (d == 0) {
        s_start[256] = sg.len;
        s_jump[0][256] = s_jump[1][256] = 256;
        s_mark[256] = 0;
    }
    block_barrier();
    const uint32_t sd = s_start[d];
    const int big = block_barrier_or(s_start[d + 1] - sd > cap);
    const uint32_t b0 = slot_begin[s], b1 = slot_begin[s + 1];
    // next(d): the first c > d with s_start[c + 1] - s_start[d] > cap, else 256
    uint32_t lo = d + 1, hi = 256;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_start[mid + 1] - sd > cap) hi = mid;
        else lo = mid + 1;
    }
    const uint32_t nx = lo;
    s_jump[0][d] = nx;
    s_mark[d] = d == 0;
    block_barrier();
    // round t: every marked node marks its 2^t-th successor; jumps double
#pragma unroll 1
    for (int t = 0; t < 8; ++t) {
        const uint32_t j = s_jump[t & 1][d];
        if (s_mark[d]) s_mark[j] = 1;  // j = 256 is the sentinel slot
        s_jump[(t & 1) ^ 1][d] = s_jump[t & 1][j];
        block_barrier();
    }
    const bool m = s_mark[d] != 0;
    const uint64_t bal = ballot(m);
    const uint32_t w = d >> 6;
    if ((d & 63) == 0) s_wave[w] = popc64(bal);
    block_barrier();
    uint32_t before = 0, ngroups = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        before += (uint32_t)i < w ? s_wave[i] : 0u;
        ngroups += s_wave[i];
    }
    const bool msd = enable && !big && ngroups <= b1 - b0;
    const uint32_t rt = msd ? kRouteMsd : kRouteLsd;
    if (d == 0) route[s] = rt;
    if (msd && m) groups[b0 + before + mbcnt(bal)] = SegSpan{sg.base + sd, s_start[nx] - sd, 0};
    for (uint32_t i = (msd ? b0 + ngroups : b0) + d; i < b1; i += 256) groups[i] = SegSpan{0, 0, 0};
    const uint32_t nt = (uint32_t)(((uint64_t)sg.len + tile_len - 1) / tile_len);
    for (uint32_t i = d; i < nt; i += 256) tiles[sg.idx + i].route = rt;
}

// odd pass count (8-bit keys): copy the large segments back from the temporaries
template <typename KB, int VB>
__global__ void k_seg_copyback(KB *keys, const KB *ktmp, typename ValT<VB>::type *vals,
                               const typename ValT<VB>::type *vtmp, const SegSpan *segs, const uint32_t *route)
{
    if (route[blockIdx.x] != kRouteLsd) return;
    const SegSpan s = segs[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < s.len; i += blockDim.x) {
        keys[s.base + i] = ktmp[s.base + i];
        if constexpr (VB > 0) vals[s.base + i] = vtmp[s.base + i];
    }
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace bcmi

// A plan: host-side routing of the segments, uploaded once.
struct bcmi_segplan_s {
    int device = 0;
    int key_dtype = 0;
    size_t value_bytes = 0;
    int descending = 0;
    uint64_t total = 0;  // elements covered by the offsets
    int small_cls = 4;      // largest small class: capacity kSegSmallThreads << small_cls
    int msd = 1;            // large segments may take the MSD route
    int ipt_large = 0;      // keys per thread of a large segment's onesweep tile
    int qrank = 0;          // 1: the large passes rank by quarter-waves (k_qsweep_seg)
    uint32_t nslots = 0;     // group slots of the MSD route, over all large segments
    uint32_t nsmall[5] = {0, 0, 0, 0, 0};  // classes IPT 1, 2, 4, 8, 16
    uint32_t ninsert = 0;
    uint32_t nlarge = 0, ntiles = 0, nchunks = 0;
    // device memory (one allocation)
    void *mem = nullptr;
    bcmi::SegSpan *small[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bcmi::SegSpan *insert = nullptr;
    bcmi::SegSpan *large = nullptr;
    bcmi::SegSpan *chunks = nullptr;
    uint32_t *route = nullptr;       // per large segment: kRouteLsd / kRouteMsd (set on the device)
    uint32_t *slot_begin = nullptr;  // per large segment: first group slot (nlarge + 1 entries)
    bcmi::SegSpan *groups = nullptr; // MSD route groups (written on the device)
    uint32_t *chunk_begin = nullptr;
    bcmi::SegTile *tiles = nullptr;
    uint32_t *partial = nullptr;
    uint32_t *seg_off = nullptr;
    void *ktmp = nullptr;
    void *vtmp = nullptr;
    // the small segments are sorted on a stream of the plan's own,
    // concurrently with the large path (the two touch disjoint elements):
    // forked from and joined back into the queue's stream, so a graph
    // capture records two branches
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

namespace bcmi {

// Keys per thread of the large segments' onesweep tiles (256 threads): 8,
// 12 or 16 by element width; for 8-byte keys or values the tuning cache
// (key "__bcmi_segsort", param "ipt") may choose 24 (twice the tile, half
// the tiles).
static int large_ipt(int kb, size_t vb)
{
    size_t sb = (size_t)kb > vb ? (size_t)kb : vb;
    if (sb == 8 && tune_get("__bcmi_segsort", "ipt", 12) == 24) return 24;
    return sb >= 16 ? 8 : (sb >= 8 ? 12 : 16);
}

// The small classes (sorted in place in LDS) and the insertion-sorted
// segments, on stream ss.
template <typename KB, int VB, typename XF>
static int run_plan_small(bcmi_queue q, bcmi_segplan pl, KB *keys, typename ValT<VB>::type *vals, XF xf,
                          hipStream_t ss)
{
#define BCMI_SMALL(CLS, IPTV)                                                                                   \
    if constexpr ((int)(sizeof(KB) + VB) * kSegSmallThreads * IPTV <= kSegSmallLds) {                           \
        if (pl->nsmall[CLS])                                                                                    \
            BCMI_LAUNCH(q, "segsort_small",                                                                     \
                        hipLaunchKernelGGL((k_seg_small<KB, VB, IPTV, XF>), dim3(pl->nsmall[CLS]),                  \
                                           dim3(kSegSmallThreads), 0, ss, keys, vals, keys, vals,                 \
                                           pl->small[CLS], xf));                                                \
    }
    BCMI_SMALL(0, 1)
    BCMI_SMALL(1, 2)
    BCMI_SMALL(2, 4)
    BCMI_SMALL(3, 8)
    BCMI_SMALL(4, 16)
#undef BCMI_SMALL
    if (pl->ninsert) {
        const unsigned blocks = (pl->ninsert + 63) / 64;
        if (pl->key_dtype == BCMI_F32) {
            if constexpr (sizeof(KB) == 4)
                BCMI_LAUNCH(q, "segsort_insertion",
                            hipLaunchKernelGGL((k_seg_insertion<float, VB>), dim3(blocks), dim3(64), 0, ss,
                                               (float *)keys, vals, pl->insert, pl->ninsert, pl->descending));
        } else {
            if constexpr (sizeof(KB) == 8)
                BCMI_LAUNCH(q, "segsort_insertion",
                            hipLaunchKernelGGL((k_seg_insertion<double, VB>), dim3(blocks), dim3(64), 0, ss,
                                               (double *)keys, vals, pl->insert, pl->ninsert, pl->descending));
        }
    }
    return 0;
}

// The large segments: top-digit route decision, MSD groups, LSD passes.
template <typename KB, int VB, typename XF, int IPTL>
static int run_plan_large(bcmi_queue q, bcmi_segplan pl, KB *keys, typename ValT<VB>::type *vals, XF xf)
{
    using V = typename ValT<VB>::type;
    constexpr int P = sizeof(KB);
    if (pl->nlarge) {
        // the top digit's histogram and offsets decide each segment's route
        // (MSD needs two digits or more); the other digits are counted only
        // for segments on the LSD route
        BCMI_LAUNCH(q, "segsort_histogram",
                    hipLaunchKernelGGL((k_seg_histogram<KB, XF>), dim3(pl->nchunks), dim3(kHistThreads), 0, q->stream,
                                       keys, pl->chunks, xf, pl->partial, (uint32_t)(P - 1), (uint32_t)P,
                                       (const uint32_t *)nullptr));
        BCMI_LAUNCH(q, "segsort_digit_offsets",
                    hipLaunchKernelGGL(k_seg_digit_offsets, dim3(pl->nlarge, 1), dim3(256), 0, q->stream, pl->partial,
                                       pl->chunk_begin, (uint32_t)P, (uint32_t)(P - 1), pl->seg_off,
                                       (const uint32_t *)nullptr));
        BCMI_LAUNCH(q, "segsort_route",
                    hipLaunchKernelGGL(k_seg_route, dim3(pl->nlarge), dim3(256), 0, q->stream, pl->large, pl->seg_off,
                                       (uint32_t)P, pl->slot_begin, pl->groups, pl->route, pl->tiles,
                                       (uint32_t)(kSortThreads * IPTL), (uint32_t)(kSegSmallThreads << pl->small_cls),
                                       (uint32_t)(P >= 2 && pl->msd)));
        if (P >= 2) {
            BCMI_LAUNCH(q, "segsort_histogram",
                        hipLaunchKernelGGL((k_seg_histogram<KB, XF>), dim3(pl->nchunks), dim3(kHistThreads), 0,
                                           q->stream, keys, pl->chunks, xf, pl->partial, 0u, (uint32_t)(P - 1),
                                           (const uint32_t *)pl->route));
            BCMI_LAUNCH(q, "segsort_digit_offsets",
                        hipLaunchKernelGGL(k_seg_digit_offsets, dim3(pl->nlarge, P - 1), dim3(256), 0, q->stream,
                                           pl->partial, pl->chunk_begin, (uint32_t)P, 0u, pl->seg_off,
                                           (const uint32_t *)pl->route));
        }
        uint64_t *status;
        BCMI_TRY(status_array(q, (size_t)pl->ntiles * 256, &status));
        BCMI_TRY(begin_lookback(q, P + 1));
        // MSD route: top-digit partition into the temporaries ...
        // (param "qrank" = 1: the quarter-wave pass body)
#define BCMI_SEG_PASS(NAME, KS, KD, VS, VD, SHIFT, PASS, BACK, WANT)                                            \
        do {                                                                                                    \
            if (pl->qrank)                                                                                      \
                BCMI_LAUNCH(q, NAME,                                                                            \
                            hipLaunchKernelGGL((k_qsweep_seg<KB, VB, IPTL, XF>), dim3(pl->ntiles),              \
                                               dim3(kSortThreads), 0, q->stream, KS, KD, VS, VD, pl->tiles,     \
                                               SHIFT, PASS, (uint32_t)P, pl->seg_off, xf, q->ctrl, status,      \
                                               BACK, WANT));                                                    \
            else                                                                                                \
                BCMI_LAUNCH(q, NAME,                                                                            \
                            hipLaunchKernelGGL((k_onesweep_seg<KB, VB, IPTL, XF>), dim3(pl->ntiles),            \
                                               dim3(kSortThreads), 0, q->stream, KS, KD, VS, VD, pl->tiles,     \
                                               SHIFT, PASS, (uint32_t)P, pl->seg_off, xf, q->ctrl, status,      \
                                               BACK, WANT));                                                    \
        } while (0)
        BCMI_SEG_PASS("segsort_onesweep_msd", keys, (KB *)pl->ktmp, vals, (V *)pl->vtmp, (uint32_t)(8 * (P - 1)),
                      (uint32_t)(P - 1), (uint32_t)(P + 1), (uint32_t)kRouteMsd);
        // ... then every group sorted in LDS from the temporaries back into place
#define BCMI_GROUPS(CLS, IPTV)                                                                                  \
    if constexpr ((int)(sizeof(KB) + VB) * kSegSmallThreads * IPTV <= kSegSmallLds) {                           \
        if (pl->small_cls == CLS && pl->nslots)                                                                 \
            BCMI_LAUNCH(q, "segsort_groups",                                                                    \
                        hipLaunchKernelGGL((k_seg_small<KB, VB, IPTV, XF>), dim3(pl->nslots), dim3(kSegSmallThreads), \
                                           0, q->stream, (const KB *)pl->ktmp, (const V *)pl->vtmp, keys, vals,   \
                                           pl->groups, xf));                                                    \
    }
        BCMI_GROUPS(0, 1)
        BCMI_GROUPS(1, 2)
        BCMI_GROUPS(2, 4)
        BCMI_GROUPS(3, 8)
        BCMI_GROUPS(4, 16)
#undef BCMI_GROUPS
        // LSD route: all P passes
        KB *ks = keys, *kd = (KB *)pl->ktmp;
        V *vs = vals, *vd = (V *)pl->vtmp;
        for (int p = 0; p < P; ++p) {
            BCMI_SEG_PASS("segsort_onesweep", ks, kd, vs, vd, (uint32_t)(8 * p), (uint32_t)p, (uint32_t)(P - p),
                          (uint32_t)kRouteLsd);
            std::swap(ks, kd);
            std::swap(vs, vd);
        }
#undef BCMI_SEG_PASS
        if (ks != keys)
            BCMI_LAUNCH(q, "segsort_copyback",
                        hipLaunchKernelGGL((k_seg_copyback<KB, VB>), dim3(pl->nlarge), dim3(256), 0, q->stream, keys,
                                           ks, vals, vs, pl->large, pl->route));
    }
    return 0;
}

template <typename KB, int VB, typename XF>
static int run_plan_xf(bcmi_queue q, bcmi_segplan pl, void *keys_, void *vals_, XF xf)
{
    using V = typename ValT<VB>::type;
    KB *keys = (KB *)keys_;
    V *vals = (V *)vals_;
    // the small segments go to the plan's stream when there is large work
    // to overlap them with (not while the queue profiles kernels: its
    // per-kernel times are taken on the queue's stream)
    bool any_small = pl->ninsert != 0;
    for (int c = 0; c < 5; ++c) any_small = any_small || pl->nsmall[c] != 0;
    const bool forked = any_small && pl->nlarge != 0 && pl->aux != nullptr && !q->profile;
    if (forked) {
        BCMI_HIP(hipEventRecord(pl->fork, q->stream));
        BCMI_HIP(hipStreamWaitEvent(pl->aux, pl->fork, 0));
    }
    int st = run_plan_small<KB, VB, XF>(q, pl, keys, vals, xf, forked ? pl->aux : q->stream);
    // the join is recorded whatever happened, so that a capture is never
    // left with a branch outside the queue's stream
    if (forked) {
        const hipError_t e = hipEventRecord(pl->join, pl->aux);
        if (e != hipSuccess && !st) st = hip_fail(e, "segmented sort: join record", __FILE__, __LINE__);
    }
    if (!st) {
        constexpr int SB = (int)sizeof(KB) > VB ? (int)sizeof(KB) : VB;
        constexpr int IPTL = SB >= 16 ? 8 : (SB >= 8 ? 12 : 16);
        // the tile size the plan's tables were built for
        bool ran = false;
        if constexpr (SB == 8) {
            if (pl->ipt_large == 24) {
                st = run_plan_large<KB, VB, XF, 24>(q, pl, keys, vals, xf);
                ran = true;
            }
        }
        if (!ran) st = run_plan_large<KB, VB, XF, IPTL>(q, pl, keys, vals, xf);
    }
    if (forked) {
        const hipError_t e = hipStreamWaitEvent(q->stream, pl->join, 0);
        if (e != hipSuccess && !st) st = hip_fail(e, "segmented sort: join", __FILE__, __LINE__);
    }
    return st;
}

template <typename KB, int VB>
static int run_plan(bcmi_queue q, bcmi_segplan pl, void *keys, void *vals)
{
    const int kind = pl->key_dtype == BCMI_F32 || pl->key_dtype == BCMI_F64 ? 2
                   : (pl->key_dtype == BCMI_I8 || pl->key_dtype == BCMI_I16 || pl->key_dtype == BCMI_I32 ||
                      pl->key_dtype == BCMI_I64) ? 1 : 0;
    return with_key_xf<KB>(kind, pl->descending,
                           [&](auto xf) { return run_plan_xf<KB, VB>(q, pl, keys, vals, xf); });
}

template <typename KB>
static int run_plan_vb(bcmi_queue q, bcmi_segplan pl, void *keys, void *vals)
{
    switch (vals ? pl->value_bytes : 0) {
    case 0: return run_plan<KB, 0>(q, pl, keys, nullptr);
    case 1: return run_plan<KB, 1>(q, pl, keys, vals);
    case 2: return run_plan<KB, 2>(q, pl, keys, vals);
    case 4: return run_plan<KB, 4>(q, pl, keys, vals);
    case 8: return run_plan<KB, 8>(q, pl, keys, vals);
    case 16: return run_plan<KB, 16>(q, pl, keys, vals);
    default: set_error("segmented sort: value size %zu unsupported", pl->value_bytes); return BCMI_INVALID_VALUE;
    }
}

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_segmented_plan_create(bcmi_queue q, int key_dtype, size_t value_bytes,
                                          const uint64_t *host_offsets, size_t nseg, int descending,
                                          bcmi_segplan *plan)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(plan && (host_offsets || nseg == 0), BCMI_INVALID_VALUE, "segmented plan: null argument");
    BCMI_CHECK(dtype_valid(key_dtype), BCMI_INVALID_VALUE, "segmented plan: bad key dtype %d", key_dtype);
    BCMI_CHECK(value_bytes == 0 || value_bytes == 1 || value_bytes == 2 || value_bytes == 4 || value_bytes == 8 ||
                   value_bytes == 16,
               BCMI_INVALID_VALUE, "segmented plan: value size %zu unsupported", value_bytes);
    const int kb = dtype_size(key_dtype);
    const bool flt = key_dtype == BCMI_F32 || key_dtype == BCMI_F64;
    const int ipt_large = large_ipt(kb, value_bytes);
    const uint64_t tile = (uint64_t)kSortThreads * ipt_large;
    SegTables tb;
    std::string err;
    // tuning cache key "__bcmi_segsort": "small_max" caps the small classes
    // (segments above go to the large path and its MSD groups of the same
    // capacity), "msd" = 0 sends every large segment down the LSD passes
    const int small_max = (int)tune_get("__bcmi_segsort", "small_max", 4);
    const int st0 = build_seg_tables(host_offsets, nseg, (size_t)kb + value_bytes, flt, tile, tb, err, small_max);
    BCMI_CHECK(st0 == 0, st0, "%s", err.c_str());
    const int cmax = tb.small_cls;
    std::vector<SegSpan>(&small)[5] = tb.small;
    std::vector<SegSpan> &insert = tb.insert, &large = tb.large, &chunks = tb.chunks;
    std::vector<uint32_t> &chunk_begin = tb.chunk_begin, &slot_begin = tb.slot_begin;
    std::vector<SegTile> &tiles = tb.tiles;
    const uint64_t total = tb.total;
    bcmi_segplan pl = new bcmi_segplan_s();
    pl->device = q->device;
    pl->key_dtype = key_dtype;
    pl->value_bytes = value_bytes;
    pl->descending = descending ? 1 : 0;
    pl->msd = tune_get("__bcmi_segsort", "msd", 1) ? 1 : 0;
    pl->ipt_large = ipt_large;
    pl->qrank = tune_get("__bcmi_segsort", "qrank", 0) ? 1 : 0;
    pl->total = total;
    pl->nlarge = (uint32_t)large.size();
    pl->ntiles = (uint32_t)tiles.size();
    pl->nchunks = (uint32_t)chunks.size();
    pl->ninsert = (uint32_t)insert.size();
    // one allocation: tables, histogram partials, digit offsets, temporaries
    size_t off = 0, o_small[5], o_insert, o_large, o_chunks, o_cb, o_tiles, o_partial, o_segoff, o_k, o_v;
    for (int c = 0; c < 5; ++c) {
        pl->nsmall[c] = (uint32_t)small[c].size();
        o_small[c] = off;
        off += al256(small[c].size() * sizeof(SegSpan));
    }
    o_insert = off; off += al256(insert.size() * sizeof(SegSpan));
    o_large = off; off += al256(large.size() * sizeof(SegSpan));
    o_chunks = off; off += al256(chunks.size() * sizeof(SegSpan));
    o_cb = off; off += al256(chunk_begin.size() * sizeof(uint32_t));
    o_tiles = off; off += al256(tiles.size() * sizeof(SegTile));
    o_partial = off; off += al256(chunks.size() * (size_t)kb * 256 * 4);
    o_segoff = off; off += al256(large.size() * (size_t)kb * 256 * 4);
    const size_t o_route = off; off += al256(large.size() * sizeof(uint32_t));
    const size_t o_sb = off; off += al256(slot_begin.size() * sizeof(uint32_t));
    const size_t o_groups = off; off += al256((size_t)slot_begin.back() * sizeof(SegSpan));
    // temporaries cover the elements of large segments at their own offsets
    const uint64_t span = large.empty() ? 0 : (large.back().base + large.back().len);
    o_k = off; off += al256(span * (size_t)kb);
    o_v = off; off += al256(span * value_bytes);
    hipError_t e = hipMalloc(&pl->mem, off + 256);
    if (e != hipSuccess) {
        delete pl;
        return hip_fail(e, "segmented plan allocation", __FILE__, __LINE__);
    }
    char *m = (char *)pl->mem;
    for (int c = 0; c < 5; ++c) pl->small[c] = (SegSpan *)(m + o_small[c]);
    pl->insert = (SegSpan *)(m + o_insert);
    pl->large = (SegSpan *)(m + o_large);
    pl->chunks = (SegSpan *)(m + o_chunks);
    pl->chunk_begin = (uint32_t *)(m + o_cb);
    pl->tiles = (SegTile *)(m + o_tiles);
    pl->partial = (uint32_t *)(m + o_partial);
    pl->seg_off = (uint32_t *)(m + o_segoff);
    pl->route = (uint32_t *)(m + o_route);
    pl->slot_begin = (uint32_t *)(m + o_sb);
    pl->groups = (SegSpan *)(m + o_groups);
    pl->nslots = slot_begin.back();
    pl->small_cls = cmax;
    pl->ktmp = m + o_k;
    pl->vtmp = m + o_v;
    auto up = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes) BCMI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    int st = 0;
    for (int c = 0; c < 5 && !st; ++c) st = up(pl->small[c], small[c].data(), small[c].size() * sizeof(SegSpan));
    if (!st) st = up(pl->insert, insert.data(), insert.size() * sizeof(SegSpan));
    if (!st) st = up(pl->large, large.data(), large.size() * sizeof(SegSpan));
    if (!st) st = up(pl->chunks, chunks.data(), chunks.size() * sizeof(SegSpan));
    if (!st) st = up(pl->chunk_begin, chunk_begin.data(), chunk_begin.size() * sizeof(uint32_t));
    if (!st) st = up(pl->tiles, tiles.data(), tiles.size() * sizeof(SegTile));
    if (!st) st = up(pl->slot_begin, slot_begin.data(), slot_begin.size() * sizeof(uint32_t));
    if (st) {
        (void)hipFree(pl->mem);
        delete pl;
        return st;
    }
    // look-back words for the large tiles, so that running the plan allocates nothing
    uint64_t *status;
    if (pl->ntiles) {
        int s2 = status_array(q, (size_t)pl->ntiles * 256, &status);
        if (s2) {
            (void)hipFree(pl->mem);
            delete pl;
            return s2;
        }
    }
    // the second stream (default priority) and its fork / join events, when
    // there are both small and large segments
    bool any_small = pl->ninsert != 0;
    for (int c = 0; c < 5; ++c) any_small = any_small || pl->nsmall[c] != 0;
    if (any_small && pl->nlarge) {
        hipError_t e2 = hipStreamCreateWithFlags(&pl->aux, hipStreamNonBlocking);
        if (e2 == hipSuccess) e2 = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming);
        if (e2 == hipSuccess) e2 = hipEventCreateWithFlags(&pl->join, hipEventDisableTiming);
        if (e2 != hipSuccess) {
            bcmi_segmented_plan_destroy(pl);
            return hip_fail(e2, "segmented plan stream", __FILE__, __LINE__);
        }
    }
    *plan = pl;
    return 0;
}

extern "C" int bcmi_segmented_plan_destroy(bcmi_segplan plan)
{
    if (!plan) return 0;
    (void)hipSetDevice(plan->device);
    if (plan->join) (void)hipEventDestroy(plan->join);
    if (plan->fork) (void)hipEventDestroy(plan->fork);
    if (plan->aux) (void)hipStreamDestroy(plan->aux);
    if (plan->mem) BCMI_HIP(hipFree(plan->mem));
    delete plan;
    return 0;
}

extern "C" int bcmi_segmented_sort_plan(bcmi_queue q, bcmi_segplan plan, void *keys, void *values)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(plan, BCMI_INVALID_VALUE, "segmented sort: null plan");
    BCMI_CHECK(plan->device == q->device, BCMI_INVALID_COMMAND_QUEUE, "segmented sort: plan is for another device");
    if (plan->total == 0) return 0;
    BCMI_CHECK(keys, BCMI_INVALID_VALUE, "segmented sort: null keys");
    if (!values) values = nullptr;
    switch (dtype_size(plan->key_dtype)) {
    case 1: return run_plan_vb<uint8_t>(q, plan, keys, values);
    case 2: return run_plan_vb<uint16_t>(q, plan, keys, values);
    case 4: return run_plan_vb<uint32_t>(q, plan, keys, values);
    default: return run_plan_vb<uint64_t>(q, plan, keys, values);
    }
}

extern "C" int bcmi_segmented_sort(bcmi_queue q, int key_dtype, void *keys, void *values, size_t value_bytes,
                                   const uint64_t *host_offsets, size_t nseg, int descending)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(!q->capturing, BCMI_INVALID_OPERATION,
               "segmented sort: build a plan (bcmi_segmented_plan_create) before capturing a graph");
    bcmi_segplan plan = nullptr;
    BCMI_TRY(bcmi_segmented_plan_create(q, key_dtype, values ? value_bytes : 0, host_offsets, nseg, descending, &plan));
    int st = bcmi_segmented_sort_plan(q, plan, keys, values);
    hipError_t e = hipStreamSynchronize(q->stream);
    bcmi_segmented_plan_destroy(plan);
    if (st) return st;
    if (e != hipSuccess) return hip_fail(e, "segmented sort", __FILE__, __LINE__);
    return 0;
}
