// scan.hip -- exclusive_scan / inclusive_scan / partial_sum.
//
// Reference: scan_on_gpu (algorithm/detail/scan_on_gpu.hpp:187-324): a
// 256-wide Hillis-Steele block scan that writes block sums, a recursive scan
// of the block sums, then a second kernel that adds them back (about 16
// bytes of traffic per element over 4+ launches; in-place scans first copy
// the input, :292-303).
//
// Here: one launch, single pass, 8 bytes per element.  Block b scans tile b
// in registers (per-thread
// chunk folds, DPP wave scans, an LDS step across the block's 4 waves), and
// obtains the tile's exclusive prefix by decoupled look-back over the
// predecessors' published aggregates.  Integer scans wrap exactly.  Float
// plus carries compensated FP32 pairs inside the tile and an FP64 carry
// between tiles.  In-place works because a tile reads its inputs before it
// writes.
#include "device.hpp"
#include "ops.hpp"

#include <cstring>

namespace bcmi {


// Integer / plain policy.  Narrow types compute in 32 bits (DPP moves 32 or
// 64 bits) and truncate on output, which keeps modular results exact.
template <typename T, int OP>
struct ScanPlain {
    using S = std::conditional_t<(sizeof(T) < 4), std::conditional_t<std::is_signed<T>::value, int32_t, uint32_t>, T>;
    using Cr = S;
    static BHD S id() { return op_identity<S, OP>(); }
    static BDEV S add(S s, T x) { return apply_op<S, OP>(s, (S)x); }
    static BDEV S merge(S a, S b) { return apply_op<S, OP>(a, b); }
    static BDEV Cr carry(S s) { return s; }
    static BHD Cr cid() { return id(); }
    static BDEV Cr cmerge(Cr a, Cr b) { return apply_op<S, OP>(a, b); }
    static BDEV T out(Cr prefix, S s) { return (T)apply_op<S, OP>(prefix, s); }
    static Cr host_init(const void *p)
    {
        T v;
        memcpy(&v, p, sizeof(T));
        return (Cr)v;
    }
};

// Compensated float plus: TwoSum pairs within the tile, FP64 between tiles.
struct ScanCompF32 {
    using S = cf32;
    using Cr = double;
    static BHD S id() { return cf32{0.f, 0.f}; }
    static BDEV S add(S s, float x) { return cf_add(s, x); }
    static BDEV S merge(S a, S b) { return cf_merge(a, b); }
    static BDEV Cr carry(S s) { return cf_value(s); }
    static BHD Cr cid() { return 0.0; }
    static BDEV Cr cmerge(Cr a, Cr b) { return a + b; }
    static BDEV float out(Cr prefix, S s) { return (float)(prefix + cf_value(s)); }
    static Cr host_init(const void *p)
    {
        float v;
        memcpy(&v, p, sizeof(float));
        return (double)v;
    }
};

// The same precision class with FP64 accumulators: each float element is
// widened once and the tile's running sums, wave scans and carries are
// doubles (53 bits instead of a TwoSum pair's 48, one add instead of six
// instructions).  FP64 vector math runs at half the FP32 rate here, so this
// is about a third of the pair arithmetic.
struct ScanF64AccF32 {
    using S = double;
    using Cr = double;
    static BHD S id() { return 0.0; }
    static BDEV S add(S s, float x) { return s + (double)x; }
    static BDEV S merge(S a, S b) { return a + b; }
    static BDEV Cr carry(S s) { return s; }
    static BHD Cr cid() { return 0.0; }
    static BDEV Cr cmerge(Cr a, Cr b) { return a + b; }
    static BDEV float out(Cr prefix, S s) { return (float)(prefix + s); }
    static Cr host_init(const void *p)
    {
        float v;
        memcpy(&v, p, sizeof(float));
        return (double)v;
    }
};

// Where a tile's elements come from.  ArraySrc: an array (every plain
// scan).  IterIn (ops.hpp): a sequence as the reference's iterators give
// it, mapped element by element.
template <typename T>
struct ArraySrc {
    const T *ptr;
    BDEV bool array() const { return true; }
    BDEV T gen(size_t) const { return T(); }
    BDEV T map(T x) const { return x; }
};

template <typename T>
struct alignas(16) sc_chunk {
    static constexpr int N = 16 / sizeof(T);
    T v[N];
};

template <typename P, int NW>
struct ScanLds {
    typename P::S wave[NW];
    typename P::Cr prefix;
};

// Striped chunks <-> blocked rows (lane l holds the wave's positions
// l*E .. l*E+E-1, E = Q*4) through LDS, G chunks at a time: the chunks of
// group g hold positions g*G*256 .. (g+1)*G*256-1, which are the rows of
// lanes g*R .. g*R+R-1 (R = G*256/E), so the wave's buffer holds R rows of
// E words, each padded by 4 words so that the 16-byte accesses of both
// directions spread over the banks.  G = Q is one group of 64 rows; G =
// Q/2 halves the buffer at the price of a second round, whose row reads
// and writes run on half the lanes.  The buffer is the wave's own: LDS
// operations of one wave execute in order, and wave_lds_sync keeps the
// compiler from moving them across each other.
template <int Q, int G>
struct RowShape {
    static constexpr int E = Q * 4, ROW = E + 4, R = G * 256 / E;
    static_assert(Q % G == 0 && R * E == G * 256 && 64 % R == 0, "groups of whole rows");
    static constexpr int WORDS = R * ROW;  // per wave
};

template <int Q, int G, typename T>
BDEV void rows_from_chunks(const T (&s)[Q][4], T (&b)[Q * 4], uint32_t *buf, uint32_t lane)
{
    using RS = RowShape<Q, G>;
    constexpr int E = RS::E, ROW = RS::ROW, R = RS::R;
#pragma unroll
    for (int g = 0; g < Q / G; ++g) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int p = j * 256 + (int)lane * 4;  // first of the chunk's 4 positions within the group
            uint4 c;
            c.x = __builtin_bit_cast(uint32_t, s[g * G + j][0]);
            c.y = __builtin_bit_cast(uint32_t, s[g * G + j][1]);
            c.z = __builtin_bit_cast(uint32_t, s[g * G + j][2]);
            c.w = __builtin_bit_cast(uint32_t, s[g * G + j][3]);
            *(uint4 *)(buf + (p / E) * ROW + p % E) = c;
        }
        wave_lds_sync();
        if (G == Q || lane / R == (uint32_t)g) {
            const uint32_t r = lane - g * R;
#pragma unroll
            for (int e = 0; e < E; e += 4) {
                const uint4 c = *(const uint4 *)(buf + r * ROW + e);
                b[e] = __builtin_bit_cast(T, c.x);
                b[e + 1] = __builtin_bit_cast(T, c.y);
                b[e + 2] = __builtin_bit_cast(T, c.z);
                b[e + 3] = __builtin_bit_cast(T, c.w);
            }
        }
        wave_lds_sync();
    }
}

template <int Q, int G, typename T>
BDEV void chunks_from_rows(const T (&b)[Q * 4], T (&s)[Q][4], uint32_t *buf, uint32_t lane)
{
    using RS = RowShape<Q, G>;
    constexpr int E = RS::E, ROW = RS::ROW, R = RS::R;
#pragma unroll
    for (int g = 0; g < Q / G; ++g) {
        if (G == Q || lane / R == (uint32_t)g) {
            const uint32_t r = lane - g * R;
#pragma unroll
            for (int e = 0; e < E; e += 4) {
                uint4 c;
                c.x = __builtin_bit_cast(uint32_t, b[e]);
                c.y = __builtin_bit_cast(uint32_t, b[e + 1]);
                c.z = __builtin_bit_cast(uint32_t, b[e + 2]);
                c.w = __builtin_bit_cast(uint32_t, b[e + 3]);
                *(uint4 *)(buf + r * ROW + e) = c;
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int p = j * 256 + (int)lane * 4;
            const uint4 c = *(const uint4 *)(buf + (p / E) * ROW + p % E);
            s[g * G + j][0] = __builtin_bit_cast(T, c.x);
            s[g * G + j][1] = __builtin_bit_cast(T, c.y);
            s[g * G + j][2] = __builtin_bit_cast(T, c.z);
            s[g * G + j][3] = __builtin_bit_cast(T, c.w);
        }
        wave_lds_sync();
    }
}

// One tile: Q 16-byte chunks per thread (tile = NW waves * 64 lanes * Q
// chunks), wave-striped: chunk q of lane l of wave w covers elements
// w*WSEG + q*64*V + l*V .. +V-1 of the tile.  FULL: the tile lies below n, so
// no element needs a bounds test.  VEC: input and output are 16-byte
// aligned, so full chunks move as one 16-byte load / store.  BG > 0 (4-byte
// elements): the loaded chunks are transposed, BG chunks at a time, into
// blocked rows through the wave's LDS buffer (blk, RowShape<Q, BG>::WORDS
// words per wave) so that each lane scans Q*4 consecutive elements serially
// and the wave scans one fold per lane instead of one per chunk; the
// results go back to striped chunks the same way.  That trades Q wave
// scans per tile for two LDS round trips.
template <typename T, typename P, bool VEC, int Q, int NW, bool FULL, int BG, typename Src>
BDEV void scan_body(ScanLds<P, NW> &L, uint32_t *blk, const Src src, T *out, size_t n, Ctrl *ctrl,
                    uint64_t *status, typename P::Cr init, int exclusive, uint32_t epoch)
{
    using S = typename P::S;
    using Cr = typename P::Cr;
    constexpr int V = 16 / sizeof(T);
    constexpr int WSEG = Q * 64 * V;  // elements per wave
    constexpr int TILE = NW * WSEG;

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t tile = blockIdx.x;
    // wave-uniform base, so every access is one 32-bit lane offset + immediates
    const size_t wstart = (size_t)tile * TILE + (size_t)__builtin_amdgcn_readfirstlane(w) * WSEG;
    const T *win = src.ptr + wstart;
    T *wout = out + wstart;
    const int64_t rem = FULL ? (int64_t)WSEG : (int64_t)n - (int64_t)wstart;
    auto ok = [&](int q, int k) { return FULL || (int64_t)(q * 64 * V + lane * V + k) < rem; };
    // blocked position e of this lane (BLK)
    auto okb = [&](int e) { return FULL || (int64_t)(lane * Q * V + e) < rem; };
    constexpr bool BLK = BG > 0;
    uint32_t *wblk = blk + w * RowShape<Q, BLK ? BG : Q>::WORDS;

    T x[Q][V];
    if (src.array()) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const T *p = win + q * 64 * V + lane * V;
            if (VEC && FULL) {
                const sc_chunk<T> c = *(const sc_chunk<T> *)p;
#pragma unroll
                for (int k = 0; k < V; ++k) x[q][k] = src.map(c.v[k]);
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) x[q][k] = ok(q, k) ? src.map(p[k]) : T();
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int k = 0; k < V; ++k) x[q][k] = src.map(src.gen(wstart + q * 64 * V + lane * V + k));
    }
    // chunk folds, then wave scans of the folds carried across q
    auto mrg = [](S a, S b) { return P::merge(a, b); };
    S carry = P::id();
    S pre[BLK ? 1 : Q];
    T b[BLK ? Q * V : 1];
    if constexpr (BLK) {
        static_assert(V == 4, "blocked tiles of 4-byte elements");
        rows_from_chunks<Q, BG>(x, b, wblk, lane);
        S t = P::id();
#pragma unroll
        for (int e = 0; e < Q * V; ++e)
            if (okb(e)) t = P::add(t, b[e]);
        const S inc = wave_inclusive_scan(t, P::id(), mrg);
        pre[0] = wave_shr1(inc, P::id());
        carry = readlane(inc, 63);
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            S t = P::id();
#pragma unroll
            for (int k = 0; k < V; ++k)
                if (ok(q, k)) t = P::add(t, x[q][k]);
            const S inc = wave_inclusive_scan(t, P::id(), mrg);
            pre[q] = P::merge(carry, wave_shr1(inc, P::id()));
            carry = P::merge(carry, readlane(inc, 63));
        }
    }
    if (lane == 0) L.wave[w] = carry;
    block_barrier();
    S wpre = P::id();
    for (uint32_t i = 0; i < w; ++i) wpre = P::merge(wpre, L.wave[i]);
    if (w == 0) {
        S total = L.wave[0];
#pragma unroll
        for (int i = 1; i < NW; ++i) total = P::merge(total, L.wave[i]);
        const Cr prefix = lookback_scan<Cr>(status, tile, epoch, P::carry(total), init, P::cid(),
                                            [](Cr a, Cr b) { return P::cmerge(a, b); }, ctrl);
        if (lane == 0) L.prefix = prefix;
    }
    block_barrier();
    const Cr prefix = L.prefix;
    if constexpr (BLK) {
        // positions past n follow every valid one in blocked order, so the
        // running sum of a valid position never includes them.  With FP64
        // accumulators the tile's prefix is added once, at the start of the
        // lane's run, rather than to every element.
        constexpr bool SEED = std::is_same<P, ScanF64AccF32>::value;
        S s = P::merge(wpre, pre[0]);
        if constexpr (SEED) s = prefix + s;
        T y[Q * V];
#pragma unroll
        for (int e = 0; e < Q * V; ++e) {
            const S before = s;
            T xk = b[e];
            if constexpr (sizeof(S) > sizeof(T)) {
                uint32_t u = __builtin_bit_cast(uint32_t, xk);
                asm volatile("" : "+v"(u));
                xk = __builtin_bit_cast(T, u);
            }
            s = P::add(s, xk);
            if constexpr (SEED)
                y[e] = (T)(exclusive ? before : s);
            else
                y[e] = P::out(prefix, exclusive ? before : s);
        }
        chunks_from_rows<Q, BG>(y, x, wblk, lane);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            T *dst = wout + q * 64 * V + lane * V;
            if (VEC && FULL) {
                sc_chunk<T> c;
#pragma unroll
                for (int k = 0; k < V; ++k) c.v[k] = x[q][k];
                *(sc_chunk<T> *)dst = c;
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k)
                    if (ok(q, k)) dst[k] = x[q][k];
            }
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        S s = P::merge(wpre, pre[q]);
        T y[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const S before = s;
            T xk = x[q][k];
            if constexpr (sizeof(S) > sizeof(T) && sizeof(T) == 4) {
                // widened again here rather than kept widened since the fold
                // (that would double the registers the tile's elements hold)
                uint32_t u = __builtin_bit_cast(uint32_t, xk);
                asm volatile("" : "+v"(u));
                xk = __builtin_bit_cast(T, u);
            }
            s = P::add(s, xk);
            y[k] = P::out(prefix, exclusive ? before : s);
        }
        T *dst = wout + q * 64 * V + lane * V;
        if (VEC && FULL) {
            sc_chunk<T> c;
#pragma unroll
            for (int k = 0; k < V; ++k) c.v[k] = y[k];
            *(sc_chunk<T> *)dst = c;
        } else {
#pragma unroll
            for (int k = 0; k < V; ++k)
                if (ok(q, k)) dst[k] = y[k];
        }
    }
}

// NW waves per workgroup (4, or 8: a tile twice as big, half the tiles and
// look-backs).
// WPE: waves per SIMD the register allocation must allow (1: no bound).
template <typename T, typename P, bool VEC, int Q, typename Src, int NW = 4, int BG = 0, int WPE = 1>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_scan(
    const Src src, T *out, size_t n, Ctrl *ctrl, uint64_t *status, typename P::Cr init, int exclusive,
    uint32_t back)
{
    constexpr size_t TILE = (size_t)NW * Q * 64 * (16 / sizeof(T));
    __shared__ ScanLds<P, NW> lds;
    uint32_t *blk = nullptr;
    if constexpr (BG > 0) {
        __shared__ uint32_t rows[NW * RowShape<Q, BG>::WORDS];
        blk = rows;
    }
    const uint32_t epoch = launch_epoch_plain(ctrl, back);
    if ((size_t)(blockIdx.x + 1) * TILE <= n)
        scan_body<T, P, VEC, Q, NW, true, BG>(lds, blk, src, out, n, ctrl, status, init, exclusive, epoch);
    else
        scan_body<T, P, VEC, Q, NW, false, BG>(lds, blk, src, out, n, ctrl, status, init, exclusive, epoch);
}

template <typename T, typename P, int Q, typename Src, int NW = 4, int BG = 0, int WPE = 1>
static int launch_scan(bcmi_queue q, const Src &src, T *out, size_t n, uint64_t *status, typename P::Cr ini,
                       int exclusive)
{
    constexpr size_t TILE = (size_t)NW * Q * 64 * (16 / sizeof(T));
    size_t tiles = (n + TILE - 1) / TILE;
    BCMI_CHECK(tiles < (1ull << 31), BCMI_INVALID_BUFFER_SIZE, "scan: n=%zu too large", n);
    bool vec = (((uintptr_t)src.ptr | (uintptr_t)out) & 15) == 0;
    BCMI_TRY(begin_lookback(q, 1));
    if (vec) {
        BCMI_LAUNCH(q, "scan",
                    hipLaunchKernelGGL((k_scan<T, P, true, Q, Src, NW, BG, WPE>), dim3((unsigned)tiles), dim3(NW * 64), 0,
                                       q->stream, src, out, n, q->ctrl, status, ini, exclusive, 1u));
    } else {
        BCMI_LAUNCH(q, "scan",
                    hipLaunchKernelGGL((k_scan<T, P, false, Q, Src, NW, BG, WPE>), dim3((unsigned)tiles), dim3(NW * 64), 0,
                                       q->stream, src, out, n, q->ctrl, status, ini, exclusive, 1u));
    }
    return 0;
}

// Chunks per thread come from the tuning cache (key "__bcmi_scan", or
// "__bcmi_scan_f32" for the compensated float scan; param "q": 4, 8 or 16;
// param "waves": 4, or 8 with q = 8 or 16) for 4-byte array scans; 4
// chunks and 4 waves otherwise.  The compensated float scan also reads
// param "blocked": lanes scan consecutive elements (BG above); 1 = one
// transpose round (4 waves and q 8 or 16, or 8 waves and q 8), 2 = two
// rounds of half the chunks each, in half the LDS, with registers bounded
// for 4 waves per SIMD (4 or 8 waves, q 8 or 16).
template <typename P>
constexpr bool comp_policy()
{
    return std::is_same<P, ScanCompF32>::value || std::is_same<P, ScanF64AccF32>::value;
}

template <typename T, typename P, typename Src>
static int run_scan_src(bcmi_queue q, const Src &src, void *out_, size_t n, const void *init, int exclusive)
{
    T *out = (T *)out_;
    const size_t tiles_max = (n + (size_t)4 * 4 * 64 - 1) / ((size_t)4 * 4 * 64) + 1;
    uint64_t *status;
    BCMI_TRY(status_array(q, tiles_max * (sizeof(typename P::Cr) / 4), &status));
    typename P::Cr ini = P::cid();
    if (exclusive && init) ini = P::host_init(init);
    if constexpr (sizeof(T) == 4 && std::is_same<Src, ArraySrc<T>>::value) {
        // the compensated float scan holds a TwoSum pair per chunk, so it has
        // its own chunk count (fewer chunks: fewer VGPRs, more resident tiles)
        const bool comp = comp_policy<P>();
        const char *key = comp ? "__bcmi_scan_f32" : "__bcmi_scan";
        const unsigned qq = tune_get(key, "q", 16);
        if constexpr (comp_policy<P>()) {
            const unsigned blocked = tune_get(key, "blocked", 0);
            const bool w8 = tune_get(key, "waves", 4) == 8;
            if (blocked == 1) {
                if (w8 && qq == 8) return launch_scan<T, P, 8, Src, 8, 8>(q, src, out, n, status, ini, exclusive);
                if (qq == 16) return launch_scan<T, P, 16, Src, 4, 16>(q, src, out, n, status, ini, exclusive);
                if (qq == 8) return launch_scan<T, P, 8, Src, 4, 8>(q, src, out, n, status, ini, exclusive);
            }
            if (blocked == 2) {
                if (w8 && qq == 16)
                    return launch_scan<T, P, 16, Src, 8, 8, 4>(q, src, out, n, status, ini, exclusive);
                if (w8 && qq == 8) return launch_scan<T, P, 8, Src, 8, 4, 4>(q, src, out, n, status, ini, exclusive);
                if (qq == 16) return launch_scan<T, P, 16, Src, 4, 8, 4>(q, src, out, n, status, ini, exclusive);
                if (qq == 8) return launch_scan<T, P, 8, Src, 4, 4, 4>(q, src, out, n, status, ini, exclusive);
            }
        }
        if (tune_get(key, "waves", 4) == 8) {
            if (qq == 16) return launch_scan<T, P, 16, Src, 8>(q, src, out, n, status, ini, exclusive);
            if (qq == 8) return launch_scan<T, P, 8, Src, 8>(q, src, out, n, status, ini, exclusive);
        }
        if (qq == 16) return launch_scan<T, P, 16>(q, src, out, n, status, ini, exclusive);
        if (qq == 8) return launch_scan<T, P, 8>(q, src, out, n, status, ini, exclusive);
    }
    return launch_scan<T, P, 4>(q, src, out, n, status, ini, exclusive);
}

template <typename T, typename P>
static int run_scan(bcmi_queue q, const void *in, void *out, size_t n, const void *init, int exclusive)
{
    return run_scan_src<T, P>(q, ArraySrc<T>{(const T *)in}, out, n, init, exclusive);
}

template <typename T>
static int scan_plain(bcmi_queue q, int op, const void *in, void *out, size_t n, const void *init, int excl)
{
    switch (op) {
    case BCMI_PLUS: return run_scan<T, ScanPlain<T, BCMI_PLUS>>(q, in, out, n, init, excl);
    case BCMI_MULTIPLIES: return run_scan<T, ScanPlain<T, BCMI_MULTIPLIES>>(q, in, out, n, init, excl);
    case BCMI_MIN: return run_scan<T, ScanPlain<T, BCMI_MIN>>(q, in, out, n, init, excl);
    case BCMI_MAX: return run_scan<T, ScanPlain<T, BCMI_MAX>>(q, in, out, n, init, excl);
    default: set_error("scan: unsupported op %d", op); return BCMI_INVALID_VALUE;
    }
}

}  // namespace bcmi

using namespace bcmi;

extern "C" int bcmi_scan(bcmi_queue q, int dtype, int op, const void *in, void *out, size_t n, const void *init,
                         int exclusive, int precision)
{
    BCMI_TRY(use_queue(q));
    if (n == 0) return 0;
    BCMI_CHECK(in && out, BCMI_INVALID_VALUE, "scan: null pointer");
    switch (dtype) {
    case BCMI_I8: return scan_plain<int8_t>(q, op, in, out, n, init, exclusive);
    case BCMI_U8: return scan_plain<uint8_t>(q, op, in, out, n, init, exclusive);
    case BCMI_I16: return scan_plain<int16_t>(q, op, in, out, n, init, exclusive);
    case BCMI_U16: return scan_plain<uint16_t>(q, op, in, out, n, init, exclusive);
    case BCMI_I32: return scan_plain<int32_t>(q, op, in, out, n, init, exclusive);
    case BCMI_U32: return scan_plain<uint32_t>(q, op, in, out, n, init, exclusive);
    case BCMI_I64: return scan_plain<int64_t>(q, op, in, out, n, init, exclusive);
    case BCMI_U64: return scan_plain<uint64_t>(q, op, in, out, n, init, exclusive);
    case BCMI_F32:
        if (op == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED) {
            // tuning cache key "__bcmi_scan_f32", param "acc": 1 = FP64
            // accumulators, 0 = TwoSum FP32 pairs
            if (tune_get("__bcmi_scan_f32", "acc", 1))
                return run_scan<float, ScanF64AccF32>(q, in, out, n, init, exclusive);
            return run_scan<float, ScanCompF32>(q, in, out, n, init, exclusive);
        }
        return scan_plain<float>(q, op, in, out, n, init, exclusive);
    case BCMI_F64: return scan_plain<double>(q, op, in, out, n, init, exclusive);
    default: set_error("scan: bad dtype %d", dtype); return BCMI_INVALID_VALUE;
    }
}

// ------------------------------------------------ fancy-iterator inputs
namespace bcmi {

template <typename T>
static int scan_iter_scalar(bcmi_queue q, int op, const bcmi_iter *in, void *out, size_t n, const void *init,
                            int excl, int precision)
{
    const IterIn<T> src = iter_in<T>(in);
    if constexpr (std::is_same<T, float>::value) {
        if (op == BCMI_PLUS && precision == BCMI_PRECISION_COMPENSATED)
            return run_scan_src<T, ScanCompF32>(q, src, out, n, init, excl);
    }
    switch (op) {
    case BCMI_PLUS: return run_scan_src<T, ScanPlain<T, BCMI_PLUS>>(q, src, out, n, init, excl);
    case BCMI_MULTIPLIES: return run_scan_src<T, ScanPlain<T, BCMI_MULTIPLIES>>(q, src, out, n, init, excl);
    case BCMI_MIN: return run_scan_src<T, ScanPlain<T, BCMI_MIN>>(q, src, out, n, init, excl);
    case BCMI_MAX: return run_scan_src<T, ScanPlain<T, BCMI_MAX>>(q, src, out, n, init, excl);
    default: set_error("scan: unsupported op %d", op); return BCMI_INVALID_VALUE;
    }
}

// vector elements: plus, component by component, from an array
template <typename E, int W>
static int scan_iter_vector(bcmi_queue q, const bcmi_iter *in, void *out, size_t n, const void *init, int excl)
{
    using T = vecn<E, W>;
    return run_scan<T, ScanPlain<T, BCMI_PLUS>>(q, in->ptr, out, n, init, excl);
}

}  // namespace bcmi

extern "C" int bcmi_scan_iter(bcmi_queue q, int dtype, int width, int op, const bcmi_iter *in, void *out, size_t n,
                              const void *init, int exclusive, int precision)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(in, BCMI_INVALID_VALUE, "scan: null input descriptor");
    BCMI_CHECK(in->kind >= BCMI_ITER_BUFFER && in->kind <= BCMI_ITER_CONSTANT, BCMI_INVALID_VALUE,
               "scan: bad iterator kind %d", in->kind);
    BCMI_CHECK((in->fn >= BCMI_MAP_ID && in->fn <= BCMI_MAP_SQUARE) ||
                   (in->fn >= BCMI_MAP_CMP && in->fn <= BCMI_MAP_CMP + BCMI_CMP_NE),
               BCMI_INVALID_VALUE, "scan: bad element map %d", in->fn);
    if (n == 0) return 0;
    BCMI_CHECK(out && (in->kind != BCMI_ITER_BUFFER || in->ptr), BCMI_INVALID_VALUE, "scan: null pointer");
    if (width == 2 || width == 4) {
        BCMI_CHECK(op == BCMI_PLUS && in->kind == BCMI_ITER_BUFFER && in->fn == BCMI_MAP_ID && in->offset == 0,
                   BCMI_INVALID_VALUE, "scan: vector elements take plus over an array only");
        switch (dtype) {
        case BCMI_I32: return width == 2 ? scan_iter_vector<int32_t, 2>(q, in, out, n, init, exclusive)
                                         : scan_iter_vector<int32_t, 4>(q, in, out, n, init, exclusive);
        case BCMI_U32: return width == 2 ? scan_iter_vector<uint32_t, 2>(q, in, out, n, init, exclusive)
                                         : scan_iter_vector<uint32_t, 4>(q, in, out, n, init, exclusive);
        case BCMI_F32: return width == 2 ? scan_iter_vector<float, 2>(q, in, out, n, init, exclusive)
                                         : scan_iter_vector<float, 4>(q, in, out, n, init, exclusive);
        default: set_error("scan: vector elements of dtype %d unsupported (int, uint, float)", dtype);
                 return BCMI_INVALID_VALUE;
        }
    }
    BCMI_CHECK(width == 1, BCMI_INVALID_VALUE, "scan: element width %d unsupported (1, 2 or 4)", width);
    // a plain array: the tuned array scan
    if (in->kind == BCMI_ITER_BUFFER && in->fn == BCMI_MAP_ID && in->offset == 0)
        return bcmi_scan(q, dtype, op, in->ptr, out, n, init, exclusive, precision);
    switch (dtype) {
    case BCMI_I32: return scan_iter_scalar<int32_t>(q, op, in, out, n, init, exclusive, precision);
    case BCMI_U32: return scan_iter_scalar<uint32_t>(q, op, in, out, n, init, exclusive, precision);
    case BCMI_I64: return scan_iter_scalar<int64_t>(q, op, in, out, n, init, exclusive, precision);
    case BCMI_U64: return scan_iter_scalar<uint64_t>(q, op, in, out, n, init, exclusive, precision);
    case BCMI_F32: return scan_iter_scalar<float>(q, op, in, out, n, init, exclusive, precision);
    case BCMI_F64: return scan_iter_scalar<double>(q, op, in, out, n, init, exclusive, precision);
    default: set_error("scan: generated or mapped inputs of dtype %d unsupported (32/64-bit int, float, double)",
                       dtype);
             return BCMI_INVALID_VALUE;
    }
}
