// select.hip -- copy_if / remove_if / count_if / unique / partition
// (stream compaction).
//
// Reference: transform_if_impl (algorithm/transform_if.hpp:33-82) writes a
// 0/1 flag per element, runs exclusive_scan over the flags, reads the count
// back (blocking) and scatters the selected elements through the scanned
// indices: 4 + ~16 + 4 + 4 bytes of traffic per element plus the output,
// over 5+ launches.  unique_copy (unique_copy.hpp:99-134) adds a transform
// for adjacent-pair flags and a gather; stable_partition
// (stable_partition.hpp) copies the input and runs copy_if twice; count_if
// (count_if.hpp, detail/count_if_with_reduce.hpp) reduces the predicate.
//
// Here: one launch reads each element once.  Block b loads tile b as
// wave-striped 16-byte chunks, evaluates the predicate into one flag word
// per thread, ranks the selected elements inside the tile (popcounts, DPP
// wave scans of two chunks' counts packed per 32-bit word, an LDS step
// across the 4 waves), obtains the number selected before the tile by
// decoupled look-back, stages the tile's output in LDS in input order and
// writes it as one contiguous run.  The output order is the input order,
// as the reference's scan-based scatter gives.  Tiles are laid over the
// 16-byte-aligned address space: positions before `in` (a buffer_iterator
// offset) are masked off, so every full chunk is one 16-byte load.
#include "device.hpp"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace bcmi {

constexpr int kSelThreads = 256;

// Comparison type: C's integer promotion makes 8- and 16-bit integers int.
template <typename T>
struct SelCmp {
    using type = T;
};
template <>
struct SelCmp<int8_t> {
    using type = int32_t;
};
template <>
struct SelCmp<uint8_t> {
    using type = int32_t;
};
template <>
struct SelCmp<int16_t> {
    using type = int32_t;
};
template <>
struct SelCmp<uint16_t> {
    using type = int32_t;
};

// A predicate as the kernels see it.  The comparison's outcome is
// ((lt & mlt) | (eq & meq) | (gt & mgt)) ^ inv: a NaN operand makes lt, eq
// and gt all false, so only != (and the negated forms) hold, as in IEEE C.
struct SelPred {
    uint64_t a, c;
    uint32_t fn, mlt, meq, mgt, inv;
    uint32_t pow2;  // a is a positive power of two (fast %)
};

template <typename C>
BHD C sel_bits(uint64_t b)
{
    C v;
    memcpy(&v, &b, sizeof(C));
    return v;
}

template <typename C>
BDEV bool sel_cmp(C y, C c, const SelPred &p)
{
    const uint32_t lt = y < c, eq = y == c, gt = y > c;
    return (((lt & p.mlt) | (eq & p.meq) | (gt & p.mgt)) ^ p.inv) != 0u;
}

// C's % (truncating; x % -1 == 0 without the INT_MIN overflow).  Not
// inlined: the division sequences of a whole tile, interleaved by the
// scheduler, would set the register budget of every predicate's kernel.
template <typename C>
__device__ __attribute__((noinline)) C sel_mod(C y, C a, uint32_t pow2)
{
    if (pow2) {
        C r = y & (a - 1);
        if constexpr (std::is_signed<C>::value)
            if (y < 0 && r != 0) r -= a;
        return r;
    }
    if constexpr (std::is_signed<C>::value)
        if (a == (C)-1) return 0;
    return y % a;
}

template <typename T>
BDEV int32_t sel_signbit(T x)
{
    if constexpr (sizeof(T) == 4) return (int32_t)(__builtin_bit_cast(uint32_t, x) >> 31);
    else return (int32_t)(__builtin_bit_cast(uint64_t, x) >> 63);
}

template <typename T, int FN>
BDEV bool sel_eval(T x, const SelPred &p)
{
    using C = typename SelCmp<T>::type;
    if constexpr (FN == BCMI_FN_SIGNBIT) {
        return sel_cmp<int32_t>(sel_signbit(x), sel_bits<int32_t>(p.c), p);
    } else {
        C y = (C)x;
        if constexpr (FN == BCMI_FN_MUL) {
            if constexpr (std::is_integral<C>::value) {
                using U = std::make_unsigned_t<C>;
                y = (C)((U)y * (U)sel_bits<C>(p.a));  // wraps
            } else {
                y = y * sel_bits<C>(p.a);
            }
        } else if constexpr (FN == BCMI_FN_MOD) {
            y = sel_mod<C>(y, sel_bits<C>(p.a), p.pow2);
        } else if constexpr (FN == BCMI_FN_AND) {
            y = y & sel_bits<C>(p.a);
        }
        return sel_cmp<C>(y, sel_bits<C>(p.c), p);
    }
}

// What a select launch writes.
enum : int {
    kSelValues = 0,   // selected elements (copy_if, remove_if)
    kSelIdx32 = 1,    // their indices as uint32 (copy_index_if)
    kSelIdx64 = 2,    // ... as uint64
    kSelPart2 = 3,    // selected to out_a, the others to out_b (partition_copy)
    kSelPartRev = 4,  // selected to out_a from the front, the others from the back (stable_partition, pass 1)
    kSelUnique = 5    // first element of each run of equal elements (unique_copy)
};

#ifndef BCMI_SEL_EPT4
#define BCMI_SEL_EPT4 64  // elements per thread, types of 1, 2 and 4 bytes
#endif
#ifndef BCMI_SEL_EPT8
#define BCMI_SEL_EPT8 32  // elements per thread, 8-byte types
#endif
#ifndef BCMI_SEL_STAGE
#define BCMI_SEL_STAGE 32768  // LDS bytes of output staging per block
#endif

template <typename T>
struct SelShape {
    static constexpr int V = 16 / (int)sizeof(T);                               // elements per chunk
    static constexpr int EPT = sizeof(T) == 8 ? BCMI_SEL_EPT8 : BCMI_SEL_EPT4;  // one flag word per thread
    static constexpr int Q = EPT / V;                                           // chunks per thread
    static constexpr int WSEG = Q * 64 * V;                                     // elements per wave
    static constexpr int TILE = 4 * WSEG;
    using F = std::conditional_t<(EPT > 32), uint64_t, uint32_t>;  // flag word
};

// The tile's output is staged in LDS in ROUNDS rounds of 4 / ROUNDS waves
// each, so the staging buffer stays within BCMI_SEL_STAGE bytes.  Index
// modes stage the position inside the tile (the index is formed when it is
// written).
template <typename T, int MODE>
struct SelLds {
    using S = std::conditional_t<MODE == kSelIdx32 || MODE == kSelIdx64, uint32_t, T>;
    static constexpr size_t BYTES = (size_t)SelShape<T>::TILE * sizeof(S);
    static constexpr int ROUNDS = BYTES > 2 * BCMI_SEL_STAGE ? 4 : BYTES > BCMI_SEL_STAGE ? 2 : 1;
    S stage[SelShape<T>::TILE / ROUNDS];
    uint32_t wcnt[4];
    uint32_t prefix;
    uint32_t bad;
};

template <typename T>
struct alignas(16) sel_chunk {
    T v[16 / sizeof(T)];
};

// Flags of the elements (bit q*V + k for chunk q, element k) under the
// predicate FN; invalid elements get 0.
template <typename T, int FN, typename OK>
BDEV typename SelShape<T>::F sel_flags(const T (&x)[SelShape<T>::Q][SelShape<T>::V], const SelPred &p, OK ok)
{
    using F = typename SelShape<T>::F;
    constexpr int V = SelShape<T>::V, Q = SelShape<T>::Q;
    F f = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int k = 0; k < V; ++k)
            if (ok(q, k) && sel_eval<T, FN>(x[q][k], p)) f |= F(1) << (q * V + k);
    return f;
}

template <typename T, int MODE, bool FULL>
BDEV void select_body(SelLds<T, MODE> &L, const T *abase, size_t head, size_t n, void *out_a, void *out_b,
                      const SelPred &p, Ctrl *ctrl, uint64_t *status, uint64_t *count_out, uint32_t epoch)
{
    using Sh = SelShape<T>;
    using C = typename SelCmp<T>::type;
    using F = typename Sh::F;
    constexpr int V = Sh::V, Q = Sh::Q, WSEG = Sh::WSEG, TILE = Sh::TILE;
    constexpr int ROUNDS = SelLds<T, MODE>::ROUNDS, WPR = 4 / ROUNDS, RSEG = WPR * WSEG;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const size_t tstart = (size_t)tile * TILE;  // aligned position of the tile
    const uint32_t wu = __builtin_amdgcn_readfirstlane(w);
    const size_t wstart = tstart + (size_t)wu * WSEG;
    const size_t end = head + n;  // aligned positions [head, end) are the range
    // chunk q of lane l covers wave positions loc = q*64*V + l*V + k, aligned
    // positions wstart + loc; the range's positions in the wave: [wlo, whi)
    const T *win = abase + wstart;
    const int32_t wlo = (int32_t)std::min<int64_t>(std::max<int64_t>((int64_t)head - (int64_t)wstart, 0), WSEG);
    const int32_t whi = (int32_t)std::min<int64_t>(std::max<int64_t>((int64_t)end - (int64_t)wstart, 0), WSEG);
    auto loc = [&](int q, int k) { return (int32_t)(q * 64 * V + lane * V + k); };
    auto ok = [&](int q, int k) {
        if constexpr (FULL) return true;
        return loc(q, k) >= wlo && loc(q, k) < whi;
    };

    T x[Q][V];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const T *src = win + loc(q, 0);
        if (FULL || (loc(q, 0) >= wlo && loc(q, V - 1) < whi)) {
            const sel_chunk<T> c = *(const sel_chunk<T> *)src;
#pragma unroll
            for (int k = 0; k < V; ++k) x[q][k] = c.v[k];
        } else {
#pragma unroll
            for (int k = 0; k < V; ++k) x[q][k] = ok(q, k) ? src[k] : T(0);
        }
    }

    F f;
    if constexpr (MODE == kSelUnique) {
        // The element before the wave's first position.  In place (out ==
        // in) it is still the input value: a tile before this one writes an
        // output index at or below it only when every element up to it was
        // kept, and then writes that same element there; a later tile
        // writes only after this tile has published its count, which
        // depends on this load.
        T before_chunk = T(0);  // element before lane 0's chunk q
        if (wstart > head && wstart < end) before_chunk = abase[wstart - 1];  // waves past the range load nothing
        f = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const C lastc = (C)x[q][V - 1];
            const T from_left = (T)dpp<0x138>(lastc, lastc);  // lane l-1's last element of chunk q
            const T prev0 = lane == 0 ? before_chunk : from_left;
            before_chunk = (T)readlane(lastc, 63);
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const T pv = k == 0 ? prev0 : x[q][k - 1];
                const bool first = loc(q, k) == wlo && wstart + wlo == head;
                if (ok(q, k) && (first || !(x[q][k] == pv))) f |= F(1) << (q * V + k);
            }
        }
    } else if constexpr (std::is_floating_point<T>::value) {
        switch (p.fn) {
        case BCMI_FN_MUL: f = sel_flags<T, BCMI_FN_MUL>(x, p, ok); break;
        case BCMI_FN_SIGNBIT: f = sel_flags<T, BCMI_FN_SIGNBIT>(x, p, ok); break;
        default: f = sel_flags<T, BCMI_FN_ID>(x, p, ok); break;
        }
    } else {
        switch (p.fn) {
        case BCMI_FN_MUL: f = sel_flags<T, BCMI_FN_MUL>(x, p, ok); break;
        case BCMI_FN_MOD: f = sel_flags<T, BCMI_FN_MOD>(x, p, ok); break;
        case BCMI_FN_AND: f = sel_flags<T, BCMI_FN_AND>(x, p, ok); break;
        default: f = sel_flags<T, BCMI_FN_ID>(x, p, ok); break;
        }
    }

    // ranks inside the wave: two chunks' counts per 32-bit word (16-bit
    // fields: a wave's count per chunk is at most 64 * 16)
    constexpr uint32_t CM = (1u << V) - 1u;
    uint32_t lrank[Q];  // selected elements before lane's chunk q, within the wave
    uint32_t running = 0;
#pragma unroll
    for (int q = 0; q < Q; q += 2) {
        const uint32_t c0 = __builtin_popcount((uint32_t)(f >> (q * V)) & CM);
        const uint32_t c1 = q + 1 < Q ? __builtin_popcount((uint32_t)(f >> ((q + 1) * V)) & CM) : 0u;
        const uint32_t packed = c0 | (c1 << 16);
        const uint32_t inc = wave_inclusive_scan(packed, 0u, [](uint32_t a, uint32_t b) { return a + b; });
        const uint32_t exc = inc - packed;
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        lrank[q] = running + (exc & 0xffffu);
        running += tot & 0xffffu;
        if (q + 1 < Q) {
            lrank[q + 1] = running + (exc >> 16);
            running += tot >> 16;
        }
    }
    if (lane == 0) L.wcnt[w] = running;
    block_barrier();
    uint32_t wpre = 0;
    for (uint32_t i = 0; i < w; ++i) wpre += L.wcnt[i];
    const uint32_t cnt = L.wcnt[0] + L.wcnt[1] + L.wcnt[2] + L.wcnt[3];
    // valid positions of the tile: [vlo, vhi) relative to tstart
    const size_t vlo_abs = tstart > head ? tstart : head;
    const size_t vhi_abs = tstart + TILE < end ? tstart + TILE : end;
    const uint32_t vcount = (uint32_t)(vhi_abs - vlo_abs);
    const uint32_t vlo = (uint32_t)(vlo_abs - tstart);
    if (w == 0) {
        const uint32_t prefix = lookback_scan<uint32_t>(status, tile, epoch, cnt, 0u, 0u,
                                                        [](uint32_t a, uint32_t b) { return a + b; }, ctrl);
        if (lane == 0) {
            L.prefix = prefix;
            // the tile's output must stay inside [0, n): a protocol error
            // drops the tile's writes and records a fault
            const size_t before = vlo_abs - head;  // valid elements before the tile
            L.bad = ((size_t)prefix + cnt > n || (size_t)prefix > before) ? 1u : 0u;
            if (L.bad) __hip_atomic_store(&ctrl->fault, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (tile == gridDim.x - 1) *count_out = (uint64_t)prefix + cnt;
        }
    }
    // stage the output in LDS in input order, one round of WPR waves at a
    // time, and write each round as contiguous runs
    constexpr bool PART = MODE == kSelPart2 || MODE == kSelPartRev;
#pragma unroll
    for (int h = 0; h < ROUNDS; ++h) {
        uint32_t rh = 0, ch = 0;  // selected before the round / in the round
        for (int i = 0; i < 4; ++i) {
            if (i < h * WPR) rh += L.wcnt[i];
            else if (i < (h + 1) * WPR) ch += L.wcnt[i];
        }
        // valid positions before the round and in it, relative to vlo
        const int64_t lo = (int64_t)h * RSEG - vlo, hi = (int64_t)(h + 1) * RSEG - vlo;
        const uint32_t vb = (uint32_t)std::min<int64_t>(std::max<int64_t>(lo, 0), vcount);
        const uint32_t vh = (uint32_t)std::min<int64_t>(std::max<int64_t>(hi, 0), vcount) - vb;
        const uint32_t uh = vb - rh;  // unselected before the round
        if ((int)(w / WPR) == h) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                uint32_t r = wpre + lrank[q];
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const bool sel = (f >> (q * V + k)) & 1u;
                    const uint32_t e = wu * WSEG + (uint32_t)loc(q, k);  // position inside the tile
                    if constexpr (PART) {
                        if (ok(q, k)) L.stage[sel ? r - rh : ch + (e - vlo - r) - uh] = x[q][k];
                    } else if (sel) {
                        if constexpr (MODE == kSelIdx32 || MODE == kSelIdx64) L.stage[r - rh] = e;
                        else L.stage[r - rh] = x[q][k];
                    }
                    r += sel ? 1u : 0u;
                }
            }
        }
        block_barrier();
        if (L.bad) return;
        const uint32_t prefix = L.prefix + rh;  // output index of the round's first selected element
        if constexpr (PART) {
            const size_t fbefore = (vlo_abs - head) - L.prefix + uh;  // unselected elements before the round
            T *ot = (T *)out_a;
            T *of = (T *)out_b;
            for (uint32_t i = tid; i < vh; i += kSelThreads) {
                const T v = L.stage[i];
                if (i < ch) ot[(size_t)prefix + i] = v;
                else if constexpr (MODE == kSelPart2) of[fbefore + (i - ch)] = v;
                else ot[n - 1 - (fbefore + (i - ch))] = v;
            }
        } else if constexpr (MODE == kSelIdx32 || MODE == kSelIdx64) {
            using I = std::conditional_t<MODE == kSelIdx32, uint32_t, uint64_t>;
            I *o = (I *)out_a;
            const size_t ibase = tstart - head;  // index of tile position 0
            for (uint32_t i = tid; i < ch; i += kSelThreads) o[(size_t)prefix + i] = (I)(ibase + L.stage[i]);
        } else {
            T *o = (T *)out_a;
            for (uint32_t i = tid; i < ch; i += kSelThreads) o[(size_t)prefix + i] = L.stage[i];
        }
        if (h + 1 < ROUNDS) block_barrier();
    }
}

#ifndef BCMI_SEL_WPE
#define BCMI_SEL_WPE 4  // minimum waves per SIMD the register allocation must allow
#endif

template <typename T, int MODE>
__global__ __launch_bounds__(kSelThreads) __attribute__((amdgpu_waves_per_eu(BCMI_SEL_WPE))) void k_select(const T *abase, size_t head, size_t n, void *out_a,
                                                        void *out_b, SelPred p, Ctrl *ctrl, uint64_t *status,
                                                        uint64_t *count_out, uint32_t back)
{
    constexpr size_t TILE = SelShape<T>::TILE;
    __shared__ SelLds<T, MODE> lds;
    const size_t tstart = (size_t)blockIdx.x * TILE;
    const uint32_t epoch = launch_epoch_plain(ctrl, back);
    if (tstart >= head && tstart + TILE <= head + n)
        select_body<T, MODE, true>(lds, abase, head, n, out_a, out_b, p, ctrl, status, count_out, epoch);
    else
        select_body<T, MODE, false>(lds, abase, head, n, out_a, out_b, p, ctrl, status, count_out, epoch);
}

// stable_partition, pass 2: out[i] = tmp[i] below the selected count, and
// the unselected elements (stored back to front) in input order after it
template <typename T>
__global__ __launch_bounds__(kSelThreads) void k_partition_unreverse(const T *tmp, T *out, size_t n,
                                                                     const uint64_t *count)
{
    const size_t nt = *count;
    const size_t stride = (size_t)gridDim.x * kSelThreads;
    for (size_t i = (size_t)blockIdx.x * kSelThreads + threadIdx.x; i < n; i += stride)
        out[i] = i < nt ? tmp[i] : tmp[n - 1 - (i - nt)];
}

__global__ void k_set_count(uint64_t *count, uint64_t v)
{
    if (threadIdx.x == 0) *count = v;
}

// count_if: a grid of a few blocks per CU streams 16-byte chunks; block
// partials, then one block sums them in a fixed order.
template <typename T, int FN>
BDEV uint32_t count_range(const T *in, size_t head, size_t n, size_t nchunks, const SelPred &p)
{
    constexpr int V = 16 / (int)sizeof(T);
    uint32_t c = 0;
    const size_t tail0 = head + nchunks * V;
    if (blockIdx.x == 0) {
        for (size_t i = threadIdx.x; i < head; i += kSelThreads) c += sel_eval<T, FN>(in[i], p);
        for (size_t i = tail0 + threadIdx.x; i < n; i += kSelThreads) c += sel_eval<T, FN>(in[i], p);
    }
    const sel_chunk<T> *body = (const sel_chunk<T> *)(in + head);
    const size_t stride = (size_t)gridDim.x * kSelThreads;
    size_t i = (size_t)blockIdx.x * kSelThreads + threadIdx.x;
    for (; i + stride < nchunks; i += 2 * stride) {
        const sel_chunk<T> a = body[i], b = body[i + stride];
#pragma unroll
        for (int k = 0; k < V; ++k) c += sel_eval<T, FN>(a.v[k], p) + sel_eval<T, FN>(b.v[k], p);
    }
    for (; i < nchunks; i += stride) {
        const sel_chunk<T> a = body[i];
#pragma unroll
        for (int k = 0; k < V; ++k) c += sel_eval<T, FN>(a.v[k], p);
    }
    return c;
}

template <typename T>
__global__ __launch_bounds__(kSelThreads) void k_count_if(const T *in, size_t head, size_t n, size_t nchunks,
                                                          SelPred p, uint64_t *partials)
{
    __shared__ uint32_t s_w[kSelThreads / 64];
    uint32_t c;
    if constexpr (std::is_floating_point<T>::value) {
        switch (p.fn) {
        case BCMI_FN_MUL: c = count_range<T, BCMI_FN_MUL>(in, head, n, nchunks, p); break;
        case BCMI_FN_SIGNBIT: c = count_range<T, BCMI_FN_SIGNBIT>(in, head, n, nchunks, p); break;
        default: c = count_range<T, BCMI_FN_ID>(in, head, n, nchunks, p); break;
        }
    } else {
        switch (p.fn) {
        case BCMI_FN_MUL: c = count_range<T, BCMI_FN_MUL>(in, head, n, nchunks, p); break;
        case BCMI_FN_MOD: c = count_range<T, BCMI_FN_MOD>(in, head, n, nchunks, p); break;
        case BCMI_FN_AND: c = count_range<T, BCMI_FN_AND>(in, head, n, nchunks, p); break;
        default: c = count_range<T, BCMI_FN_ID>(in, head, n, nchunks, p); break;
        }
    }
    c = wave_reduce(c, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    block_barrier();
    if (threadIdx.x == 0) partials[blockIdx.x] = (uint64_t)s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(kSelThreads) void k_count_final(const uint64_t *partials, uint32_t nparts,
                                                             uint64_t *count)
{
    __shared__ uint64_t s[kSelThreads];
    uint64_t c = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += kSelThreads) c += partials[i];
    s[threadIdx.x] = c;
    block_barrier();
    for (int off = kSelThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        block_barrier();
    }
    if (threadIdx.x == 0) *count = s[0];
}

// ---------------------------------------------------------------- host side

static bool dtype_is_float(int dt) { return dt == BCMI_F32 || dt == BCMI_F64; }

// Host view of a predicate: validated, comparison masks precomputed.
static int make_pred(int dtype, const bcmi_predicate *pr, SelPred *p)
{
    BCMI_CHECK(pr, BCMI_INVALID_VALUE, "predicate is null");
    BCMI_CHECK(pr->fn >= BCMI_FN_ID && pr->fn <= BCMI_FN_SIGNBIT, BCMI_INVALID_VALUE, "predicate: bad fn %d", pr->fn);
    BCMI_CHECK(pr->cmp >= BCMI_CMP_LT && pr->cmp <= BCMI_CMP_NE, BCMI_INVALID_VALUE, "predicate: bad cmp %d",
               pr->cmp);
    const bool fl = dtype_is_float(dtype);
    BCMI_CHECK(!(fl && (pr->fn == BCMI_FN_MOD || pr->fn == BCMI_FN_AND)), BCMI_INVALID_VALUE,
               "predicate: %% and & need an integer element type");
    BCMI_CHECK(!(!fl && pr->fn == BCMI_FN_SIGNBIT), BCMI_INVALID_VALUE, "predicate: signbit needs a float type");
    memset(p, 0, sizeof(*p));
    p->a = pr->operand;
    p->c = pr->value;
    p->fn = (uint32_t)pr->fn;
    static const uint32_t masks[6][4] = {
        // lt, eq, gt, inv
        {1, 0, 0, 0}, {1, 1, 0, 0}, {0, 0, 1, 0}, {0, 1, 1, 0}, {0, 1, 0, 0}, {0, 1, 0, 1}};
    p->mlt = masks[pr->cmp][0];
    p->meq = masks[pr->cmp][1];
    p->mgt = masks[pr->cmp][2];
    p->inv = masks[pr->cmp][3] ^ (pr->negate ? 1u : 0u);
    if (pr->fn == BCMI_FN_MOD) {
        // the operand in the comparison type
        int64_t s = 0;
        uint64_t u = 0;
        bool is_signed = true;
        switch (dtype) {
        case BCMI_I8: case BCMI_U8: case BCMI_I16: case BCMI_U16: case BCMI_I32:
            s = sel_bits<int32_t>(pr->operand); break;
        case BCMI_U32: u = sel_bits<uint32_t>(pr->operand); is_signed = false; break;
        case BCMI_I64: s = sel_bits<int64_t>(pr->operand); break;
        default: u = pr->operand; is_signed = false; break;
        }
        BCMI_CHECK(is_signed ? s != 0 : u != 0, BCMI_INVALID_VALUE, "predicate: %% by zero");
        if (is_signed) p->pow2 = s > 0 && (s & (s - 1)) == 0;
        else p->pow2 = (u & (u - 1)) == 0;
    }
    return 0;
}

template <typename T, int MODE>
static int launch_select(bcmi_queue q, const T *in, void *out_a, void *out_b, size_t n, const SelPred &p,
                         uint64_t *count_out)
{
    constexpr size_t TILE = SelShape<T>::TILE;
    const size_t head = ((uintptr_t)in & 15) / sizeof(T);
    BCMI_CHECK(((uintptr_t)in % sizeof(T)) == 0, BCMI_INVALID_VALUE, "select: input not element-aligned");
    const T *abase = in - head;
    const size_t tiles = (n + head + TILE - 1) / TILE;
    BCMI_CHECK(tiles < (1ull << 31), BCMI_INVALID_BUFFER_SIZE, "select: n=%zu too large", n);
    uint64_t *status;
    BCMI_TRY(status_array(q, tiles + 1, &status));
    BCMI_TRY(begin_lookback(q, 1));
    BCMI_LAUNCH(q, "select",
                hipLaunchKernelGGL((k_select<T, MODE>), dim3((unsigned)tiles), dim3(kSelThreads), 0, q->stream,
                                   abase, head, n, out_a, out_b, p, q->ctrl, status, count_out, 1u));
    return 0;
}

template <typename T>
static int run_copy_if(bcmi_queue q, const void *in, void *out, size_t n, const SelPred &p, int kind,
                       uint64_t *count)
{
    switch (kind) {
    case BCMI_SELECT_VALUES: return launch_select<T, kSelValues>(q, (const T *)in, out, nullptr, n, p, count);
    case BCMI_SELECT_INDEX_U32: return launch_select<T, kSelIdx32>(q, (const T *)in, out, nullptr, n, p, count);
    case BCMI_SELECT_INDEX_U64: return launch_select<T, kSelIdx64>(q, (const T *)in, out, nullptr, n, p, count);
    default: set_error("copy_if: bad out_kind %d", kind); return BCMI_INVALID_VALUE;
    }
}

template <typename T>
static int run_unique(bcmi_queue q, const void *in, void *out, size_t n, uint64_t *count)
{
    SelPred p;
    memset(&p, 0, sizeof(p));
    return launch_select<T, kSelUnique>(q, (const T *)in, out, nullptr, n, p, count);
}

template <typename T>
static int run_partition(bcmi_queue q, const void *in, void *out_t, void *out_f, size_t n, const SelPred &p,
                         uint64_t *count)
{
    if (out_f) return launch_select<T, kSelPart2>(q, (const T *)in, out_t, out_f, n, p, count);
    // both groups into out_t: through a temporary (out_t may be in)
    void *tmp;
    BCMI_TRY(scratch(q, n * sizeof(T), &tmp));
    BCMI_TRY((launch_select<T, kSelPartRev>(q, (const T *)in, tmp, nullptr, n, p, count)));
    const unsigned blocks = (unsigned)std::min<size_t>((n + kSelThreads - 1) / kSelThreads, (size_t)cu_blocks(q, 8));
    BCMI_LAUNCH(q, "partition_unreverse",
                hipLaunchKernelGGL((k_partition_unreverse<T>), dim3(blocks), dim3(kSelThreads), 0, q->stream,
                                   (const T *)tmp, (T *)out_t, n, count));
    return 0;
}

template <typename T>
static int run_count_if(bcmi_queue q, const void *in_, size_t n, const SelPred &p, uint64_t *count)
{
    const T *in = (const T *)in_;
    constexpr int V = 16 / (int)sizeof(T);
    size_t head = 0;
    const uintptr_t mis = (uintptr_t)in & 15;
    if (mis) head = (mis % sizeof(T)) ? n : (16 - mis) / sizeof(T);
    if (head > n) head = n;
    const size_t nchunks = (n - head) / V;
    const size_t want = (nchunks + kSelThreads - 1) / kSelThreads;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(want, (size_t)cu_blocks(q, 4)));
    void *sp;
    BCMI_TRY(scratch(q, blocks * sizeof(uint64_t), &sp));
    uint64_t *partials = (uint64_t *)sp;
    BCMI_LAUNCH(q, "count_if",
                hipLaunchKernelGGL((k_count_if<T>), dim3(blocks), dim3(kSelThreads), 0, q->stream, in, head, n,
                                   nchunks, p, partials));
    BCMI_LAUNCH(q, "count_final",
                hipLaunchKernelGGL(k_count_final, dim3(1), dim3(kSelThreads), 0, q->stream, (const uint64_t *)partials,
                                   blocks, count));
    return 0;
}

static int set_count(bcmi_queue q, uint64_t *count, uint64_t v)
{
    BCMI_LAUNCH(q, "set_count", hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, q->stream, count, v));
    return 0;
}

#define BCMI_SEL_DISPATCH(dtype, FN, ...)                                    \
    switch (dtype) {                                                         \
    case BCMI_I8: return FN<int8_t>(__VA_ARGS__);                            \
    case BCMI_U8: return FN<uint8_t>(__VA_ARGS__);                           \
    case BCMI_I16: return FN<int16_t>(__VA_ARGS__);                          \
    case BCMI_U16: return FN<uint16_t>(__VA_ARGS__);                         \
    case BCMI_I32: return FN<int32_t>(__VA_ARGS__);                          \
    case BCMI_U32: return FN<uint32_t>(__VA_ARGS__);                         \
    case BCMI_I64: return FN<int64_t>(__VA_ARGS__);                          \
    case BCMI_U64: return FN<uint64_t>(__VA_ARGS__);                         \
    case BCMI_F32: return FN<float>(__VA_ARGS__);                            \
    case BCMI_F64: return FN<double>(__VA_ARGS__);                           \
    default: set_error("bad dtype %d", dtype); return BCMI_INVALID_VALUE;    \
    }

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_copy_if(bcmi_queue q, int dtype, const void *in, void *out, size_t n, const bcmi_predicate *pred,
                            int out_kind, uint64_t *count_out)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(count_out, BCMI_INVALID_VALUE, "copy_if: count_out is null");
    SelPred p;
    BCMI_TRY(make_pred(dtype, pred, &p));
    BCMI_CHECK(dtype_valid(dtype), BCMI_INVALID_VALUE, "copy_if: bad dtype %d", dtype);
    if (n == 0) return set_count(q, count_out, 0);
    BCMI_CHECK(in && out, BCMI_INVALID_VALUE, "copy_if: null pointer");
    BCMI_CHECK(n < (1ull << 32), BCMI_INVALID_BUFFER_SIZE, "copy_if: n=%zu must be below 2^32", n);
    BCMI_CHECK(out_kind == BCMI_SELECT_VALUES || in != out, BCMI_INVALID_VALUE,
               "copy_if: indices cannot be written over the input");
    BCMI_SEL_DISPATCH(dtype, run_copy_if, q, in, out, n, p, out_kind, count_out)
}

extern "C" int bcmi_count_if(bcmi_queue q, int dtype, const void *in, size_t n, const bcmi_predicate *pred,
                             uint64_t *count_out)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(count_out, BCMI_INVALID_VALUE, "count_if: count_out is null");
    SelPred p;
    BCMI_TRY(make_pred(dtype, pred, &p));
    BCMI_CHECK(dtype_valid(dtype), BCMI_INVALID_VALUE, "count_if: bad dtype %d", dtype);
    if (n == 0) return set_count(q, count_out, 0);
    BCMI_CHECK(in, BCMI_INVALID_VALUE, "count_if: null pointer");
    BCMI_SEL_DISPATCH(dtype, run_count_if, q, in, n, p, count_out)
}

extern "C" int bcmi_unique(bcmi_queue q, int dtype, const void *in, void *out, size_t n, uint64_t *count_out)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(count_out, BCMI_INVALID_VALUE, "unique: count_out is null");
    BCMI_CHECK(dtype_valid(dtype), BCMI_INVALID_VALUE, "unique: bad dtype %d", dtype);
    if (n == 0) return set_count(q, count_out, 0);
    BCMI_CHECK(in && out, BCMI_INVALID_VALUE, "unique: null pointer");
    BCMI_CHECK(n < (1ull << 32), BCMI_INVALID_BUFFER_SIZE, "unique: n=%zu must be below 2^32", n);
    BCMI_SEL_DISPATCH(dtype, run_unique, q, in, out, n, count_out)
}

extern "C" int bcmi_partition(bcmi_queue q, int dtype, const void *in, void *out_true, void *out_false, size_t n,
                              const bcmi_predicate *pred, uint64_t *count_out)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(count_out, BCMI_INVALID_VALUE, "partition: count_out is null");
    SelPred p;
    BCMI_TRY(make_pred(dtype, pred, &p));
    BCMI_CHECK(dtype_valid(dtype), BCMI_INVALID_VALUE, "partition: bad dtype %d", dtype);
    if (n == 0) return set_count(q, count_out, 0);
    BCMI_CHECK(in && out_true, BCMI_INVALID_VALUE, "partition: null pointer");
    BCMI_CHECK(n < (1ull << 32), BCMI_INVALID_BUFFER_SIZE, "partition: n=%zu must be below 2^32", n);
    BCMI_CHECK(!out_false || (out_false != in && out_true != in), BCMI_INVALID_VALUE,
               "partition_copy: outputs must not be the input");
    BCMI_SEL_DISPATCH(dtype, run_partition, q, in, out_true, out_false, n, p, count_out)
}
