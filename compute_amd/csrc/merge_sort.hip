// merge_sort.hip -- comparison sort for user comparators (sort, sort_by_key,
// stable_sort, stable_sort_by_key with a comparator that is not less<T> /
// greater<T> on a fundamental key).
//
// Reference: merge_sort_on_gpu (algorithm/detail/merge_sort_on_gpu.hpp):
// a bitonic / odd-even block sort in local memory (:85-298, not stable;
// the stable variant skips it, :309-311) followed by log2(n / block) merge
// passes in which every work-item binary-searches its own element's rank in
// the partner block (:321-465), evaluating the user's OpenCL comparator on
// whole elements in global memory at every probe.
//
// Here the comparator is one of the forms of bcmi_comparator: key(a) CMP
// key(b) over fields of the records.  Each element's key is computed once,
// when the block sort loads it, into an order-preserving unsigned code
// (32-bit for integer and float fields of up to 4 bytes, 64-bit otherwise;
// greater inverts the code), so every later comparison is one unsigned
// compare of codes carried in an 8- or 16-byte item {code, index}.  A
// comparator whose two sides read different fields (a.s0 < b.s3) carries
// both codes (24-byte item) and compares code_l(a) < code_r(b).
//
//   k_ms_block   256 threads x IPT items: load (coalesced, through LDS),
//                codes, a stable odd-even transposition sort of each
//                thread's IPT items in registers, then log2(256) merge rounds
//                in LDS, each thread placing its IPT outputs by a merge-path
//                search; writes sorted runs of TILE items.
//   k_ms_split   the merge-path split at the first output of every merge
//                tile, one wave per tile: a 64-ary search (one ballot per
//                round over 64 probes), all tiles at once.
//   k_ms_merge   one merge pass of runs of length w: each block produces
//                TILE consecutive outputs between its split and the next
//                tile's, stages its A and B slices in LDS and merges them as
//                above.
//   emit         the last launch (the block sort when n <= TILE, else the
//                last merge pass) writes no items: it gathers each output's
//                record and value from copies of the inputs, by the item's
//                index, straight into the caller's arrays.
//
// Every merge is stable (ties take the left run first), so the result is the
// stable sort by the comparator for every strict weak order.
#include "device.hpp"

#include <algorithm>
#include <type_traits>

namespace bcmi {

constexpr int kMsThreads = 256;

struct CmpDev {
    int form;
    uint32_t lo, ro, o2, count;
    int greater;
};

// Items: an order code and the element's input index.
struct ItemS32 {
    uint32_t k;
    uint32_t idx;
};
struct alignas(16) ItemS64 {
    uint64_t k;
    uint32_t idx;
    uint32_t pad;
};
struct ItemA64 {
    uint64_t kl, kr;
    uint32_t idx;
    uint32_t pad;
};

BDEV bool item_less(const ItemS32 &a, const ItemS32 &b) { return a.k < b.k; }
BDEV bool item_less(const ItemS64 &a, const ItemS64 &b) { return a.k < b.k; }
BDEV bool item_less(const ItemA64 &a, const ItemA64 &b) { return a.kl < b.kr; }

// Field-wise loads: a whole-struct copy through a selected pointer stays a
// memcpy into private memory.
BDEV ItemS32 ld_item(const ItemS32 *p) { return ItemS32{p->k, p->idx}; }
BDEV ItemS64 ld_item(const ItemS64 *p) { return ItemS64{p->k, p->idx, 0u}; }
BDEV ItemA64 ld_item(const ItemA64 *p) { return ItemA64{p->kl, p->kr, p->idx, 0u}; }

// Padding items sort after every real item of the tile (a pad is never less
// than anything, and pads follow the real items, so stability keeps them last).
template <typename Item>
BDEV Item pad_item()
{
    Item it;
    if constexpr (std::is_same<Item, ItemA64>::value) {
        it.kl = it.kr = ~0ull;
    } else {
        it.k = ~decltype(it.k)(0);
    }
    it.idx = 0xffffffffu;
    return it;
}

template <typename T>
BDEV T ld_field(const unsigned char *p)
{
    return *(const T *)p;  // alignment checked on the host
}

// Order codes: unsigned compares of the codes give the comparator's order.
// Floats: -0 is mapped to +0 first (C compares them equal).
BDEV uint32_t f32_code(float f)
{
    if (f == 0.0f) f = 0.0f;
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

BDEV uint64_t f64_code(double d)
{
    if (d == 0.0) d = 0.0;
    uint64_t b = (uint64_t)__double_as_longlong(d);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

// key(record at `off`) as a C-bit code.
template <typename T, typename C>
BDEV C key_code(const unsigned char *rec, uint32_t off, const CmpDev &c)
{
    C code;
    if constexpr (std::is_floating_point<T>::value) {
        if (c.form == BCMI_KEY_LENGTH) {
            // sum of squares in double, in component order, no contraction
            double ss = 0.0;
            for (uint32_t k = 0; k < c.count; ++k) {
                double x = (double)ld_field<T>(rec + off + k * sizeof(T));
                double p = x * x;
                asm volatile("" : "+v"(p));
                ss += p;
            }
            code = (C)f64_code(ss);
        } else {
            T x = ld_field<T>(rec + off);
            if (c.form == BCMI_KEY_ABS_SUM2) x = x + ld_field<T>(rec + c.o2);
            if (c.form != BCMI_KEY_FIELD) x = __builtin_fabs(x);
            if constexpr (sizeof(T) == 4)
                code = (C)f32_code(x);
            else
                code = (C)f64_code(x);
        }
    } else {
        // C's promotion: 8- and 16-bit integers compute in int
        using P = typename std::conditional<(sizeof(T) < 4), int32_t, T>::type;
        using UP = typename std::make_unsigned<P>::type;
        P s = (P)ld_field<T>(rec + off);
        if (c.form == BCMI_KEY_FIELD) {
            if constexpr (std::is_signed<P>::value)
                code = (C)((UP)s ^ ((UP)1 << (8 * sizeof(P) - 1)));
            else
                code = (C)s;
        } else {
            if (c.form == BCMI_KEY_ABS_SUM2) s = (P)((UP)s + (UP)(P)ld_field<T>(rec + c.o2));
            UP m = (UP)s;
            if constexpr (std::is_signed<P>::value) {
                if (s < 0) m = (UP)0 - (UP)s;  // abs returns the unsigned magnitude
            }
            code = (C)m;
        }
    }
    return c.greater ? (C)~code : code;
}

template <typename T, typename Item>
BDEV Item make_item(const unsigned char *rec, const CmpDev &c, uint32_t idx)
{
    Item it;
    if constexpr (std::is_same<Item, ItemA64>::value) {
        it.kl = key_code<T, uint64_t>(rec, c.lo, c);
        it.kr = key_code<T, uint64_t>(rec, c.ro, c);
    } else {
        it.k = key_code<T, decltype(it.k)>(rec, c.lo, c);
    }
    it.idx = idx;
    return it;
}

// The final gather: records (and values) by the items' input indices.
struct Emit {
    const unsigned char *rsrc;
    unsigned char *rdst;
    uint32_t rb, rw;  // bytes per record, bytes per copy word
    const unsigned char *vsrc;
    unsigned char *vdst;
    uint32_t vb, vw;
};

BDEV void copy_words(unsigned char *dst, const unsigned char *src, uint32_t bytes, uint32_t w)
{
    switch (w) {
    case 16:
        for (uint32_t o = 0; o < bytes; o += 16) *(uint4 *)(dst + o) = *(const uint4 *)(src + o);
        break;
    case 8:
        for (uint32_t o = 0; o < bytes; o += 8) *(uint64_t *)(dst + o) = *(const uint64_t *)(src + o);
        break;
    case 4:
        for (uint32_t o = 0; o < bytes; o += 4) *(uint32_t *)(dst + o) = *(const uint32_t *)(src + o);
        break;
    case 2:
        for (uint32_t o = 0; o < bytes; o += 2) *(uint16_t *)(dst + o) = *(const uint16_t *)(src + o);
        break;
    default:
        for (uint32_t o = 0; o < bytes; ++o) dst[o] = src[o];
    }
}

BDEV void emit_one(const Emit &e, size_t pos, uint32_t idx)
{
    copy_words(e.rdst + pos * e.rb, e.rsrc + (size_t)idx * e.rb, e.rb, e.rw);
    if (e.vdst) copy_words(e.vdst + pos * e.vb, e.vsrc + (size_t)idx * e.vb, e.vb, e.vw);
}

// LDS position of tile item i: one padding slot after every IPT items, so
// the IPT consecutive items of consecutive threads start IPT + 1 slots apart
// and a wave's per-thread accesses spread over the banks instead of all
// landing on the same ones.
template <int IPT>
BDEV int lpos(int i)
{
    return i + i / IPT;
}

// Merge path inside LDS: the number of A items among the first d outputs of
// the stable merge of A = tile items [a0, a0 + na) and B = [b0, b0 + nb).
template <typename Item, int IPT>
BDEV int lds_path(const Item *lds, int a0, int na, int b0, int nb, int d)
{
    int lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (!item_less(lds[lpos<IPT>(b0 + d - 1 - mid)], lds[lpos<IPT>(a0 + mid)]))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Serial stable merge of IPT outputs starting at (i, j); outputs past `cnt`
// (the merged length minus d) hold a copy of A[0] and are not to be used.
template <typename Item, int IPT>
BDEV void lds_merge(const Item *lds, int a0, int na, int b0, int nb, int i, int j, int cnt, Item (&x)[IPT])
{
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const bool take_a = j >= nb || (i < na && !item_less(lds[lpos<IPT>(b0 + j)], lds[lpos<IPT>(a0 + i)]));
        const int src = take_a ? a0 + i : b0 + j;
        x[k] = ld_item(&lds[lpos<IPT>(k < cnt ? src : a0)]);  // one load, always in bounds
        i += take_a ? 1 : 0;
        j += take_a ? 0 : 1;
    }
}

template <typename T, typename Item, int IPT>
__global__ __launch_bounds__(kMsThreads) void k_ms_block(const unsigned char *__restrict__ rec, uint32_t rb,
                                                        size_t n, CmpDev c, Item *__restrict__ out, Emit e)
{
    constexpr int TILE = kMsThreads * IPT;
    __shared__ Item lds[TILE + TILE / IPT];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * TILE;

#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int i = k * kMsThreads + t;
        const size_t g = base + i;
        lds[lpos<IPT>(i)] = g < n ? make_item<T, Item>(rec + g * rb, c, (uint32_t)g) : pad_item<Item>();
    }
    block_barrier();
    Item x[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) x[k] = lds[lpos<IPT>(t * IPT + k)];

    // odd-even transposition sort: IPT rounds sort IPT items; swapping only
    // strictly smaller right neighbours keeps it stable
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
#pragma unroll
        for (int k = r & 1; k + 1 < IPT; k += 2) {
            if (item_less(x[k + 1], x[k])) {
                Item tmp = x[k];
                x[k] = x[k + 1];
                x[k + 1] = tmp;
            }
        }
    }

    for (int w = IPT; w < TILE; w *= 2) {
        block_barrier();
#pragma unroll
        for (int k = 0; k < IPT; ++k) lds[lpos<IPT>(t * IPT + k)] = x[k];
        block_barrier();
        const int o = t * IPT;
        const int p0 = o & ~(2 * w - 1);
        const int d = o - p0;
        const int i = lds_path<Item, IPT>(lds, p0, w, p0 + w, w, d);
        lds_merge<Item, IPT>(lds, p0, w, p0 + w, w, i, d - i, IPT, x);
    }

    block_barrier();
#pragma unroll
    for (int k = 0; k < IPT; ++k) lds[lpos<IPT>(t * IPT + k)] = x[k];
    block_barrier();
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int i = k * kMsThreads + t;
        const size_t g = base + i;
        if (g < n) {
            if (e.rdst)
                emit_one(e, g, lds[lpos<IPT>(i)].idx);
            else
                out[g] = lds[lpos<IPT>(i)];
        }
    }
}

// Merge path over global memory by one wave: the number of A items among
// the first d outputs of the stable merge of A[0, na) and B[0, nb).  Each
// round probes 64 evenly spaced candidates and keeps the interval between
// the last probe that belongs to the prefix and the first that does not.
template <typename Item>
BDEV size_t wave_path(const Item *A, size_t na, const Item *B, size_t nb, size_t d, int lane)
{
    size_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (hi > lo) {
        const size_t span = hi - lo;
        if (span <= 64) {
            bool p = false;
            if ((size_t)lane < span) {
                const size_t m = lo + lane;
                p = !item_less(B[d - 1 - m], A[m]);
            }
            return lo + (size_t)__popcll(__ballot(p));
        }
        const size_t m = lo + ((size_t)(lane + 1) * span) / 65;  // strictly increasing, < hi
        const bool p = !item_less(B[d - 1 - m], A[m]);
        const int cnt = __popcll(__ballot(p));
        const size_t m_lo = (size_t)__shfl((unsigned long long)m, cnt > 0 ? cnt - 1 : 0);
        const size_t m_hi = (size_t)__shfl((unsigned long long)m, cnt < 64 ? cnt : 63);
        if (cnt > 0) lo = m_lo + 1;
        if (cnt < 64) hi = m_hi;
    }
    return lo;
}

// The merge-path split at the first output of every merge tile: one wave
// per tile, all tiles at once, so the merge kernel starts streaming at once.
template <typename Item, int IPT>
__global__ __launch_bounds__(kMsThreads) void k_ms_split(const Item *__restrict__ in, size_t n, size_t w,
                                                        size_t tiles, size_t *__restrict__ splits)
{
    constexpr int TILE = kMsThreads * IPT;
    const size_t b = (size_t)blockIdx.x * (kMsThreads / 64) + (threadIdx.x >> 6);
    if (b >= tiles) return;  // whole waves
    const size_t o0 = b * TILE;
    const size_t p0 = o0 / (2 * w) * (2 * w);
    const size_t a_end = std::min(p0 + w, n), b_end = std::min(p0 + 2 * w, n);
    const size_t i = wave_path(in + p0, a_end - p0, in + a_end, b_end - a_end, o0 - p0, threadIdx.x & 63);
    if ((threadIdx.x & 63) == 0) splits[b] = i;
}

template <typename Item, int IPT>
__global__ __launch_bounds__(kMsThreads) void k_ms_merge(const Item *__restrict__ in, Item *__restrict__ out, size_t n,
                                                        size_t w, const size_t *__restrict__ splits, Emit e)
{
    constexpr int TILE = kMsThreads * IPT;
    __shared__ Item lds[TILE + TILE / IPT];
    const int t = threadIdx.x;
    const size_t o0 = (size_t)blockIdx.x * TILE;
    const size_t p0 = o0 / (2 * w) * (2 * w);
    const size_t a_end = std::min(p0 + w, n), b_end = std::min(p0 + 2 * w, n);
    const size_t na_all = a_end - p0, nb_all = b_end - a_end;
    const size_t d0 = o0 - p0, d1 = std::min(d0 + TILE, na_all + nb_all);

    // the tile ends where the next tile of the same pair starts, or at the end
    // of the pair (all of A taken)
    const bool last_in_pair = o0 + TILE >= p0 + 2 * w || o0 + TILE >= n;
    const size_t i0 = splits[blockIdx.x];
    const size_t i1_raw = last_in_pair ? na_all : splits[blockIdx.x + 1];
    // a comparator that is not a strict weak order can make the end split
    // inconsistent with the start; clamping keeps every access in bounds
    const size_t i1 = std::min(std::max(i1_raw, i0), i0 + (d1 - d0));
    const size_t j0 = d0 - i0, j1 = d1 - i1;
    const int na = (int)(i1 - i0), nb = (int)(j1 - j0), cnt = na + nb;

    for (int x = t; x < cnt; x += kMsThreads)
        lds[lpos<IPT>(x)] = ld_item(x < na ? in + (p0 + i0 + x) : in + (a_end + j0 + (x - na)));
    block_barrier();

    // threads past the end run the merge with no outputs (d = cnt)
    Item x[IPT];
    const int dl = t * IPT < cnt ? t * IPT : cnt;
    const int il = lds_path<Item, IPT>(lds, 0, na, na, nb, dl);
    lds_merge<Item, IPT>(lds, 0, na, na, nb, il, dl - il, cnt - dl, x);
    block_barrier();
#pragma unroll
    for (int k = 0; k < IPT; ++k)
        if (dl + k < cnt) lds[lpos<IPT>(dl + k)] = x[k];
    block_barrier();
    for (int y = t; y < cnt; y += kMsThreads) {
        if (e.rdst)
            emit_one(e, o0 + y, lds[lpos<IPT>(y)].idx);
        else
            out[o0 + y] = lds[lpos<IPT>(y)];
    }
}

// The key route: each record's order code and input index, then (after the
// stable radix sort of the codes) the gather of records and values.
template <typename T, typename C>
__global__ __launch_bounds__(kMsThreads) void k_ms_codes(const unsigned char *__restrict__ rec, uint32_t rb, size_t n,
                                                        CmpDev c, C *__restrict__ codes, uint32_t *__restrict__ idx)
{
    const size_t stride = (size_t)gridDim.x * kMsThreads;
    for (size_t i = (size_t)blockIdx.x * kMsThreads + threadIdx.x; i < n; i += stride) {
        codes[i] = key_code<T, C>(rec + i * rb, c.lo, c);
        idx[i] = (uint32_t)i;
    }
}

__global__ __launch_bounds__(kMsThreads) void k_ms_gather(const uint32_t *__restrict__ idx, size_t n, Emit e)
{
    const size_t stride = (size_t)gridDim.x * kMsThreads;
    for (size_t i = (size_t)blockIdx.x * kMsThreads + threadIdx.x; i < n; i += stride) emit_one(e, i, idx[i]);
}

template <typename Item>
struct MsIpt;
template <>
struct MsIpt<ItemS32> {
    static constexpr int value = 16;
};
template <>
struct MsIpt<ItemS64> {
    static constexpr int value = 8;
};
template <>
struct MsIpt<ItemA64> {
    static constexpr int value = 8;
};

// Items per thread of the merge passes (their tiles need not match the
// block sort's, as long as they divide it).  Half-size merge tiles made the
// 2^24 int2_ merges 10 % faster but doubled the split searches, a net loss,
// so they match.
template <typename Item>
struct MsMergeIpt {
    static constexpr int value = MsIpt<Item>::value;
};

static uint32_t copy_word(const void *p, size_t bytes)
{
    uint32_t w = 16;
    while (w > 1 && (bytes % w != 0 || (uintptr_t)p % w != 0)) w >>= 1;
    return w;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename T, typename Item>
static int run_merge_sort(bcmi_queue q, const CmpDev &c, unsigned char *rec, size_t rb, unsigned char *vals,
                          size_t vb, size_t n)
{
    constexpr int IPT = MsIpt<Item>::value, MIPT = MsMergeIpt<Item>::value;
    constexpr size_t TILE = (size_t)kMsThreads * IPT, MTILE = (size_t)kMsThreads * MIPT;
    const size_t blocks = (n + TILE - 1) / TILE, mblocks = (n + MTILE - 1) / MTILE;
    const size_t rbytes = align256(n * rb), vbytes = vals ? align256(n * vb) : 0;
    const size_t ibytes = n > TILE ? align256(n * sizeof(Item)) : 0;
    const size_t sbytes = n > TILE ? align256(mblocks * sizeof(size_t)) : 0;
    void *sp = nullptr;
    BCMI_TRY(scratch(q, rbytes + vbytes + 2 * ibytes + sbytes, &sp));
    unsigned char *rcopy = (unsigned char *)sp;
    unsigned char *vcopy = rcopy + rbytes;
    Item *buf[2] = {(Item *)(rcopy + rbytes + vbytes), (Item *)(rcopy + rbytes + vbytes + ibytes)};
    size_t *splits = (size_t *)(rcopy + rbytes + vbytes + 2 * ibytes);

    BCMI_HIP(hipMemcpyAsync(rcopy, rec, n * rb, hipMemcpyDeviceToDevice, q->stream));
    if (vals) BCMI_HIP(hipMemcpyAsync(vcopy, vals, n * vb, hipMemcpyDeviceToDevice, q->stream));
    Emit emit{rcopy, rec, (uint32_t)rb, std::min(copy_word(rec, rb), copy_word(rcopy, rb)),
              vals ? vcopy : nullptr, vals, (uint32_t)vb, vals ? std::min(copy_word(vals, vb), copy_word(vcopy, vb)) : 1u};
    Emit none{};

    BCMI_LAUNCH(q, "merge_sort_block",
                hipLaunchKernelGGL((k_ms_block<T, Item, IPT>), dim3((unsigned)blocks), dim3(kMsThreads), 0,
                                   q->stream, rcopy, (uint32_t)rb, n, c, buf[0], n <= TILE ? emit : none));
    int src = 0;
    for (size_t w = TILE; w < n; w *= 2) {
        const bool last = 2 * w >= n;
        const size_t split_blocks = (mblocks + kMsThreads / 64 - 1) / (kMsThreads / 64);
        BCMI_LAUNCH(q, "merge_sort_split",
                    hipLaunchKernelGGL((k_ms_split<Item, MIPT>), dim3((unsigned)split_blocks), dim3(kMsThreads), 0,
                                       q->stream, buf[src], n, w, mblocks, splits));
        BCMI_LAUNCH(q, "merge_sort_merge",
                    hipLaunchKernelGGL((k_ms_merge<Item, MIPT>), dim3((unsigned)mblocks), dim3(kMsThreads), 0,
                                       q->stream, buf[src], buf[src ^ 1], n, w, splits, last ? emit : none));
        src ^= 1;
    }
    return 0;
}

// Key route for comparators that read the same field of both operands:
// key(a) < key(b) is then a strict weak order by one code per element, so
// the stable radix sort of (code, index) pairs gives the stable comparator
// order in 4 (32-bit codes) or 8 passes instead of log2(n / 4096) merge
// passes, followed by one gather.  The radix sort takes the front of the
// queue's scratch arena; this route's copies and codes sit after it.
template <typename T, typename C>
__global__ __launch_bounds__(kMsThreads) void k_ms_codes_only(const unsigned char *__restrict__ rec, uint32_t rb,
                                                             size_t n, CmpDev c, C *__restrict__ codes)
{
    const size_t stride = (size_t)gridDim.x * kMsThreads;
    for (size_t i = (size_t)blockIdx.x * kMsThreads + threadIdx.x; i < n; i += stride)
        codes[i] = key_code<T, C>(rec + i * rb, c.lo, c);
}

template <typename T, typename C>
static int run_key_route(bcmi_queue q, const CmpDev &c, unsigned char *rec, size_t rb, unsigned char *vals,
                         size_t vb, size_t n)
{
    const unsigned blocks = (unsigned)std::min<size_t>((n + kMsThreads - 1) / kMsThreads, (size_t)cu_blocks(q, 8));
    if (!vals && rb <= 16 && (uintptr_t)rec % rb == 0) {
        // records of up to 16 bytes, aligned to their size, and no values:
        // the records themselves are the radix sort's values (moved as
        // rb-byte words), sorted in place; no copies, no gather
        const size_t front = align256(radix_scratch_bytes(n, sizeof(C), rb, q->cu_count));
        void *sp = nullptr;
        BCMI_TRY(scratch(q, front + align256(n * sizeof(C)), &sp));
        C *codes = (C *)((unsigned char *)sp + front);
        BCMI_LAUNCH(q, "merge_sort_codes",
                    hipLaunchKernelGGL((k_ms_codes_only<T, C>), dim3(blocks), dim3(kMsThreads), 0, q->stream, rec,
                                       (uint32_t)rb, n, c, codes));
        return bcmi_sort(q, BCMI_STABLE_SORT_BY_KEY, sizeof(C) == 4 ? BCMI_U32 : BCMI_U64, codes, rec, rb, n, 0);
    }
    const size_t front = align256(radix_scratch_bytes(n, sizeof(C), 4, q->cu_count));
    const size_t rbytes = align256(n * rb), vbytes = vals ? align256(n * vb) : 0;
    const size_t cbytes = align256(n * sizeof(C)), ibytes = align256(n * 4);
    void *sp = nullptr;
    BCMI_TRY(scratch(q, front + rbytes + vbytes + cbytes + ibytes, &sp));
    unsigned char *rcopy = (unsigned char *)sp + front;
    unsigned char *vcopy = rcopy + rbytes;
    C *codes = (C *)(rcopy + rbytes + vbytes);
    uint32_t *idx = (uint32_t *)(rcopy + rbytes + vbytes + cbytes);

    BCMI_HIP(hipMemcpyAsync(rcopy, rec, n * rb, hipMemcpyDeviceToDevice, q->stream));
    if (vals) BCMI_HIP(hipMemcpyAsync(vcopy, vals, n * vb, hipMemcpyDeviceToDevice, q->stream));
    BCMI_LAUNCH(q, "merge_sort_codes",
                hipLaunchKernelGGL((k_ms_codes<T, C>), dim3(blocks), dim3(kMsThreads), 0, q->stream, rcopy,
                                   (uint32_t)rb, n, c, codes, idx));
    BCMI_TRY(bcmi_sort(q, BCMI_STABLE_SORT_BY_KEY, sizeof(C) == 4 ? BCMI_U32 : BCMI_U64, codes, idx, 4, n, 0));
    Emit emit{rcopy, rec, (uint32_t)rb, std::min(copy_word(rec, rb), copy_word(rcopy, rb)),
              vals ? vcopy : nullptr, vals, (uint32_t)vb, vals ? std::min(copy_word(vals, vb), copy_word(vcopy, vb)) : 1u};
    BCMI_LAUNCH(q, "merge_sort_gather",
                hipLaunchKernelGGL(k_ms_gather, dim3(blocks), dim3(kMsThreads), 0, q->stream, idx, n, emit));
    return 0;
}

template <typename T>
static int merge_sort_typed(bcmi_queue q, const CmpDev &c, bool asym, unsigned char *rec, size_t rb,
                            unsigned char *vals, size_t vb, size_t n)
{
    if (asym) return run_merge_sort<T, ItemA64>(q, c, rec, rb, vals, vb, n);
    // route 1 (default): key-based comparators sort their codes by radix;
    // route 0: always the merge kernels.  The radix sort takes n up to
    // 2^32 - 2^20; above that the merge kernels serve every route.
    const bool key_route = tune_get("__bcmi_merge_sort", "route", 1) != 0 && n <= (1ull << 32) - (1ull << 20);
    if constexpr (sizeof(T) <= 4) {
        if (c.form != BCMI_KEY_LENGTH) {
            if (key_route) return run_key_route<T, uint32_t>(q, c, rec, rb, vals, vb, n);
            return run_merge_sort<T, ItemS32>(q, c, rec, rb, vals, vb, n);
        }
    }
    if constexpr (sizeof(T) == 8 || std::is_same<T, float>::value) {
        if (key_route) return run_key_route<T, uint64_t>(q, c, rec, rb, vals, vb, n);
        return run_merge_sort<T, ItemS64>(q, c, rec, rb, vals, vb, n);
    } else {
        set_error("merge_sort: length() needs float or double fields");
        return BCMI_INVALID_VALUE;
    }
}

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_merge_sort(bcmi_queue q, const bcmi_comparator *cmp, void *records, size_t record_bytes,
                               void *values, size_t value_bytes, size_t n)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(cmp, BCMI_INVALID_VALUE, "merge_sort: null comparator");
    BCMI_CHECK(dtype_valid(cmp->dtype), BCMI_INVALID_VALUE, "merge_sort: bad dtype %d", cmp->dtype);
    BCMI_CHECK(cmp->form >= BCMI_KEY_FIELD && cmp->form <= BCMI_KEY_LENGTH, BCMI_INVALID_VALUE,
               "merge_sort: bad key form %d", cmp->form);
    const size_t rb = record_bytes, sz = (size_t)dtype_size(cmp->dtype);
    BCMI_CHECK(rb == 1 || rb == 2 || rb == 4 || rb == 8 || rb == 16 || rb == 32 || rb == 64, BCMI_INVALID_VALUE,
               "merge_sort: record size %zu not 1, 2, 4, 8, 16, 32 or 64 bytes", rb);
    BCMI_CHECK(!values || value_bytes >= 1, BCMI_INVALID_VALUE, "merge_sort: value_bytes is 0");
    BCMI_CHECK(n < 0xffffffffull, BCMI_INVALID_VALUE, "merge_sort: n >= 2^32");
    const bool sym_only = cmp->form == BCMI_KEY_ABS_SUM2 || cmp->form == BCMI_KEY_LENGTH;
    BCMI_CHECK(!sym_only || cmp->lhs_offset == cmp->rhs_offset, BCMI_INVALID_VALUE,
               "merge_sort: abs-of-sum and length comparators read the same fields on both sides");
    BCMI_CHECK(cmp->form != BCMI_KEY_LENGTH || cmp->dtype == BCMI_F32 || cmp->dtype == BCMI_F64, BCMI_INVALID_VALUE,
               "merge_sort: length() needs float or double fields");
    BCMI_CHECK(cmp->form != BCMI_KEY_LENGTH || (cmp->count >= 1 && cmp->count <= 16), BCMI_INVALID_VALUE,
               "merge_sort: length() of %u components", cmp->count);
    // every field read lies inside the record and is aligned to its type
    size_t fields[3] = {cmp->lhs_offset, cmp->rhs_offset, cmp->lhs_offset};
    if (cmp->form == BCMI_KEY_ABS_SUM2) fields[2] = cmp->offset2;
    if (cmp->form == BCMI_KEY_LENGTH) fields[2] = cmp->lhs_offset + (size_t)(cmp->count - 1) * sz;
    for (size_t f : fields) {
        BCMI_CHECK(f + sz <= rb, BCMI_INVALID_VALUE, "merge_sort: field at offset %zu past the %zu-byte record", f,
                   rb);
        BCMI_CHECK(f % sz == 0, BCMI_INVALID_VALUE, "merge_sort: field at offset %zu not aligned to its type", f);
    }
    if (n < 2) return 0;
    BCMI_CHECK(records, BCMI_INVALID_VALUE, "merge_sort: null records");
    BCMI_CHECK((uintptr_t)records % sz == 0, BCMI_INVALID_VALUE, "merge_sort: records not aligned to the field type");

    CmpDev c{cmp->form, cmp->lhs_offset, cmp->rhs_offset, cmp->offset2, cmp->count, cmp->greater ? 1 : 0};
    const bool asym = cmp->lhs_offset != cmp->rhs_offset;
    unsigned char *r = (unsigned char *)records, *v = (unsigned char *)values;
    switch (cmp->dtype) {
    case BCMI_I8: return merge_sort_typed<int8_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_U8: return merge_sort_typed<uint8_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_I16: return merge_sort_typed<int16_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_U16: return merge_sort_typed<uint16_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_I32: return merge_sort_typed<int32_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_U32: return merge_sort_typed<uint32_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_I64: return merge_sort_typed<int64_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_U64: return merge_sort_typed<uint64_t>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_F32: return merge_sort_typed<float>(q, c, asym, r, rb, v, value_bytes, n);
    case BCMI_F64: return merge_sort_typed<double>(q, c, asym, r, rb, v, value_bytes, n);
    }
    return BCMI_INVALID_VALUE;
}
