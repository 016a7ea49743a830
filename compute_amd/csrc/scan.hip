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

// One tile: Q 16-byte chunks per thread (tile = NW waves * 64 lanes * Q
// chunks), wave-striped: chunk q of lane l of wave w covers elements
// w*WSEG + q*64*V + l*V .. +V-1 of the tile.  FULL: the tile lies below n, so
// no element needs a bounds test.  VEC: input and output are 16-byte
// aligned, so full chunks move as one 16-byte load / store.
template <typename T, typename P, bool VEC, int Q, int NW, bool FULL, typename Src>
BDEV void scan_body(ScanLds<P, NW> &L, const Src src, T *out, size_t n, Ctrl *ctrl, uint64_t *status,
                    typename P::Cr init, int exclusive, uint32_t epoch)
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
    S pre[Q];
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
template <typename T, typename P, bool VEC, int Q, typename Src, int NW = 4>
__global__ __launch_bounds__(NW * 64) void k_scan(const Src src, T *out, size_t n, Ctrl *ctrl,
                                                  uint64_t *status, typename P::Cr init, int exclusive,
                                                  uint32_t back)
{
    constexpr size_t TILE = (size_t)NW * Q * 64 * (16 / sizeof(T));
    __shared__ ScanLds<P, NW> lds;
    const uint32_t epoch = launch_epoch_plain(ctrl, back);
    if ((size_t)(blockIdx.x + 1) * TILE <= n)
        scan_body<T, P, VEC, Q, NW, true>(lds, src, out, n, ctrl, status, init, exclusive, epoch);
    else
        scan_body<T, P, VEC, Q, NW, false>(lds, src, out, n, ctrl, status, init, exclusive, epoch);
}

template <typename T, typename P, int Q, typename Src, int NW = 4>
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
                    hipLaunchKernelGGL((k_scan<T, P, true, Q, Src, NW>), dim3((unsigned)tiles), dim3(NW * 64), 0,
                                       q->stream, src, out, n, q->ctrl, status, ini, exclusive, 1u));
    } else {
        BCMI_LAUNCH(q, "scan",
                    hipLaunchKernelGGL((k_scan<T, P, false, Q, Src, NW>), dim3((unsigned)tiles), dim3(NW * 64), 0,
                                       q->stream, src, out, n, q->ctrl, status, ini, exclusive, 1u));
    }
    return 0;
}

// Chunks per thread come from the tuning cache (key "__bcmi_scan", or
// "__bcmi_scan_f32" for the compensated float scan; param "q": 4, 8 or 16;
// param "waves": 4, or 8 with q = 8 or 16) for 4-byte array scans; 4
// chunks and 4 waves otherwise.
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
        const bool comp = std::is_same<P, ScanCompF32>::value || std::is_same<P, ScanF64AccF32>::value;
        const char *key = comp ? "__bcmi_scan_f32" : "__bcmi_scan";
        const unsigned qq = tune_get(key, "q", 16);
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
