"""Pins the CPU oracle (oracle/bcmi_oracle.c) before anything is checked
against it.  Every oracle function is tied to at least one of: the C++
standard's known answers, the reference's own golden vectors
(tests/golden/reference_vectors.json), closed forms worked out by hand from
the reference's algorithm, or brute force in Python.  No GPU needed.
"""
import itertools
import math
import struct
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- RNG / KAT
def test_mt19937_known_answers():
    # C++11 [rand.predef]: the 10000th consecutive invocation of a
    # default-constructed engine (seed 5489) produces these values.
    assert int(O.mt19937(5489, 10000)[-1]) == 4123659995
    assert int(O.mt19937_64(5489, 10000)[-1]) == 9981545732273789042


def test_workload_generators_follow_their_definitions():
    raw = O.mt19937(0xB0057001, 64)
    assert np.array_equal(O.gen_sort_u32(64), raw)
    raw3 = O.mt19937(0xB0057003, 1000).astype(np.int64)
    x = O.gen_scan_i32(1000)
    assert np.array_equal(x, (raw3 % 201 - 100).astype(np.int32))
    assert x.min() >= -100 and x.max() <= 100
    k, v = O.gen_sort_by_key_f32(1 << 16)
    assert np.array_equal(v, np.arange(1 << 16, dtype=np.uint32))
    assert np.all(np.abs(k) <= 1e6)
    zeros = np.count_nonzero(k == 0) / k.size
    assert 0.007 < zeros < 0.013  # ~1% signed zeros
    assert np.count_nonzero(np.signbit(k) & (k == 0)) > 0
    off = O.gen_segments(65536)
    lens = np.diff(off.astype(np.int64))
    assert lens[0] == 1 << 20 and lens.min() >= 1 and lens.max() <= 1 << 20
    assert 200 < np.median(lens) < 320  # median 256
    rk, rv = O.gen_runs_i32_f32(1 << 16)
    assert np.all(np.diff(rk) >= 0) and np.all(np.diff(rk) <= 1)
    runs = np.diff(np.flatnonzero(np.diff(np.r_[-1, rk, rk[-1] + 1])))
    assert 25 < runs.mean() < 40  # geometric, mean 32


# ----------------------------------------------------------- radix() keys
INTS = {
    "uint8": [0, 1, 2, 127, 128, 200, 255],
    "int8": [-128, -127, -1, 0, 1, 100, 127],
    "uint16": [0, 1, 32767, 32768, 65535],
    "int16": [-32768, -1, 0, 1, 32767],
    "uint32": [0, 1, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 1],
    "int32": [-2 ** 31, -2 ** 31 + 1, -1, 0, 1, 2 ** 31 - 1],
    "uint64": [0, 1, 2 ** 63 - 1, 2 ** 63, 2 ** 64 - 1],
    "int64": [-2 ** 63, -1, 0, 1, 2 ** 63 - 1],
}


def bits_of(arr):
    a = np.ascontiguousarray(arr)
    return a.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.itemsize])


@pytest.mark.parametrize("dt", list(INTS))
def test_radix_key_orders_integers_numerically(dt):
    info = np.iinfo(dt)
    rnd = RNG.integers(info.min, info.max, 200, endpoint=True, dtype=dt)
    vals = np.r_[np.array(INTS[dt], dtype=dt), rnd]
    b = bits_of(vals)
    asc = [O.radix_key(int(x), dt) for x in b]
    order = np.argsort(np.array(asc, dtype=np.uint64), kind="stable")
    assert np.all(np.diff(vals[order].astype(np.float64)) >= 0)
    desc = [O.radix_key(int(x), dt, True) for x in b]
    order = np.argsort(np.array(desc, dtype=np.uint64), kind="stable")
    got = vals[order]
    if np.dtype(dt).kind == "i":
        # (-x) ^ signbit: INT_MIN negates to itself and maps to key 0, so it
        # comes first; everything else is in descending order.
        mn = np.iinfo(dt).min
        nmin = int(np.count_nonzero(vals == mn))
        assert np.all(got[:nmin] == mn)
        assert np.all(np.diff(got[nmin:].astype(np.float64)) <= 0)
    else:
        assert np.all(np.diff(got.astype(np.float64)) <= 0)


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_radix_key_orders_floats_by_ieee_total_order(dt):
    fi = np.finfo(dt)
    special = [-np.inf, -fi.max, -1.0, -fi.tiny, -fi.smallest_subnormal, -0.0, 0.0,
               fi.smallest_subnormal, fi.tiny, 1.0, fi.max, np.inf]
    vals = np.array(special + list(RNG.normal(size=300) * 10.0 ** RNG.integers(-30, 30, 300)),
                    dtype=dt)
    b = bits_of(vals)
    asc = np.array([O.radix_key(int(x), dt) for x in b], dtype=np.uint64)
    s = vals[np.argsort(asc, kind="stable")]
    assert np.all(np.diff(s) >= 0)
    # -0.0 sorts strictly before +0.0 (bit order)
    z = np.flatnonzero(s == 0)
    assert np.signbit(s[z[0]]) and not np.signbit(s[z[-1]])
    # NaNs: negative-signed NaN before -inf, positive NaN after +inf
    nan_bits = {"float32": (0x7FC00000, 0xFFC00000), "float64": (0x7FF8000000000000, 0xFFF8000000000000)}[dt]
    pos_nan = O.radix_key(nan_bits[0], dt)
    neg_nan = O.radix_key(nan_bits[1], dt)
    assert neg_nan < asc.min() and pos_nan > asc.max()
    # descending: numerically non-increasing, except that +-denorm_min share
    # keys with -+0.0 (see the collision test below), so they are left out
    keep = np.abs(vals) != fi.smallest_subnormal
    desc = np.array([O.radix_key(int(x), dt, True) for x in b[keep]], dtype=np.uint64)
    s = vals[keep][np.argsort(desc, kind="stable")]
    assert np.all(np.diff(s) <= 0)


def test_radix_key_descending_float_collisions():
    # (-x) ^ (-(x>>31) | signbit) maps +denorm_min and -0.0 to the same key,
    # and +0.0 and -denorm_min to the same key.
    f = lambda v: struct.unpack("<I", struct.pack("<f", v))[0]
    dmin = float(np.finfo(np.float32).smallest_subnormal)
    assert O.radix_key(f(dmin), "float32", True) == O.radix_key(f(-0.0), "float32", True) == 0x7FFFFFFF
    assert O.radix_key(f(0.0), "float32", True) == O.radix_key(f(-dmin), "float32", True) == 0x80000000


# ------------------------------------------------------ golden: sort paths
def _np(vals, dt):
    if dt.startswith("float"):
        return np.array(vals, dtype=dt)
    return np.array(vals, dtype=np.int64).astype(dt)


def _same(got, exp, cmp):
    if cmp == "eq":
        return np.array_equal(got, exp)
    return np.array_equal(bits_of(got), bits_of(exp))


def test_golden_sorts(golden):
    n = 0
    for c in golden:
        if c["algo"] not in ("radix_sort", "sort", "radix_sort_range"):
            continue
        x = _np(c["input"], c["dtype"])
        exp = _np(c["expected"], c["dtype"])
        desc = c.get("descending", False)
        if c["algo"] == "radix_sort":
            got = O.radix_order_sort(x, descending=desc)
        elif c["algo"] == "sort":
            got = O.sort_dispatch(O.SORT, x, descending=desc)
        else:
            got = x.copy()
            got[c["first"]:c["last"]] = O.radix_order_sort(x[c["first"]:c["last"]], descending=desc)
        assert _same(got, exp, c.get("cmp")), c["id"]
        n += 1
    assert n > 30


def test_golden_sort_by_key(golden):
    n = 0
    for c in golden:
        if c["algo"] not in ("radix_sort_by_key", "sort_by_key", "stable_sort_by_key_prefix"):
            continue
        k = _np(c["keys"], c["dtype"])
        v = _np(c["values"], c["vdtype"])
        desc = c.get("descending", False)
        api = {"radix_sort_by_key": None, "sort_by_key": O.SORT_BY_KEY,
               "stable_sort_by_key_prefix": O.STABLE_SORT_BY_KEY}[c["algo"]]
        if api is None:
            gk, gv = O.radix_order_sort(k, v, descending=desc)
        else:
            gk, gv = O.sort_dispatch(api, k, v, descending=desc)
        if c["algo"] == "stable_sort_by_key_prefix":
            m = len(c["expected_keys_prefix"])
            assert list(gk[:m]) == c["expected_keys_prefix"] and list(gv[:m]) == c["expected_values_prefix"]
        else:
            assert np.array_equal(gk, _np(c["expected_keys"], c["dtype"])), c["id"]
            assert np.array_equal(gv, _np(c["expected_values"], c["vdtype"])), c["id"]
        n += 1
    assert n >= 10


@pytest.mark.parametrize("dt", ["int32", "uint8", "float32", "float64", "int64"])
def test_radix_order_sort_is_the_stable_sort_by_radix_key(dt):
    # brute force: Python's sorted() is stable; key = radix key
    if dt.startswith("float"):
        k = RNG.integers(-8, 8, 500).astype(dt) * 0.5
        k[RNG.integers(0, 500, 20)] = -0.0
    else:
        info = np.iinfo(dt)
        k = RNG.integers(max(info.min, -20), min(info.max, 20), 500, endpoint=True).astype(dt)
    v = np.arange(k.size, dtype=np.uint32)
    for desc in (False, True):
        gk, gv = O.radix_order_sort(k, v, descending=desc)
        b = bits_of(k)
        ref = sorted(range(k.size), key=lambda i: O.radix_key(int(b[i]), dt, desc))
        assert np.array_equal(gv, np.array(ref, dtype=np.uint32))
        assert np.array_equal(bits_of(gk), b[ref])


@pytest.mark.parametrize("dt", ["int32", "float32", "uint16"])
def test_insertion_sort_is_a_stable_comparison_sort(dt):
    for n in (0, 1, 2, 5, 17, 31, 32):
        k = (RNG.integers(-5, 5, n) if dt != "uint16" else RNG.integers(0, 9, n)).astype(dt)
        if dt == "float32" and n:
            k[::3] = -0.0
            k[1::4] = 0.0
        v = np.arange(n, dtype=np.int32)
        for desc in (False, True):
            gk, gv = O.insertion_sort(k, v, descending=desc)
            ref = sorted(range(n), key=lambda i: float(k[i]), reverse=False)
            if desc:
                # a stable descending sort keeps equal keys in input order
                ref = sorted(range(n), key=lambda i: -float(k[i]))
            assert np.array_equal(gv, np.array(ref, dtype=np.int32)), (n, desc)


def test_sort_dispatch_thresholds():
    # +0.0 before -0.0 in the input: comparison sorts keep that order, the
    # radix path puts -0.0 first (sort.hpp:49-54: n <= 32 -> insertion).
    for n, api, radix in ((32, O.SORT, False), (33, O.SORT, True), (31, O.SORT_BY_KEY, False),
                          (32, O.SORT_BY_KEY, True), (2, O.STABLE_SORT, True),
                          (31, O.STABLE_SORT_BY_KEY, False)):
        k = np.ones(n, np.float32)
        k[0], k[1] = 0.0, -0.0
        v = np.arange(n, dtype=np.int32)
        gk, gv = O.sort_dispatch(api, k, v)
        assert (gv[0] == 1) == radix, (n, api)


def test_compare_stable_sort_treats_signed_zeros_equal():
    k = np.array([0.0, -0.0, 1.0, -0.0, 0.0], np.float32)
    v = np.arange(5, dtype=np.int32)
    gk, gv = O.compare_stable_sort(k, v)
    assert list(gv) == [0, 1, 3, 4, 2]


def test_segmented_sort_sorts_each_segment():
    off = np.array([0, 3, 3, 10, 11, 40], np.uint64)
    k = RNG.integers(0, 2 ** 63, 40, dtype=np.uint64)
    v = np.arange(40, dtype=np.uint32)
    gk, gv = O.segmented_sort(k, v, off)
    for s in range(len(off) - 1):
        b, e = int(off[s]), int(off[s + 1])
        ref = sorted(range(b, e), key=lambda i: int(k[i]))
        assert list(gv[b:e]) == ref


# ------------------------------------------------------------------- scans
def test_golden_scans(golden):
    n = 0
    for c in golden:
        if c["algo"] not in ("inclusive_scan", "exclusive_scan"):
            continue
        x = _np(c["input"], c["dtype"])
        op = {"plus": O.PLUS, "multiplies": O.MULTIPLIES}[c.get("op", "plus")]
        got = O.scan(x, c["algo"] == "exclusive_scan", c.get("init", 0), op)
        assert np.array_equal(got, _np(c["expected"], c["dtype"])), c["id"]
        n += 1
    assert n >= 7


@pytest.mark.parametrize("dt", ["int32", "uint8", "int64", "uint32"])
def test_integer_scans_wrap_like_numpy(dt):
    x = RNG.integers(-(2 ** 20), 2 ** 20, 5000).astype(dt)
    assert np.array_equal(O.scan(x, False), np.cumsum(x, dtype=dt))
    e = O.scan(x, True, 7)
    assert e[0] == 7 and np.array_equal(e[1:], (np.cumsum(x, dtype=dt) + np.array(7, dt)).astype(dt)[:-1])
    m = O.scan(x, False, op=O.MAX)
    assert np.array_equal(m, np.maximum.accumulate(x))
    m = O.scan(x, False, op=O.MIN)
    assert np.array_equal(m, np.minimum.accumulate(x))


def test_double_double_prefix_is_correctly_rounded():
    x = O.gen_lognormal_f32(3000)
    got = O.scan_f32_dd(x, False)
    for i in RNG.integers(0, 3000, 200):
        assert got[i] == math.fsum(float(t) for t in x[:i + 1])
    ex = O.scan_f32_dd(x, True, 0.5)
    assert ex[0] == 0.5 and ex[10] == math.fsum([0.5] + [float(t) for t in x[:10]])


def test_double_double_sum_is_correctly_rounded():
    for n in (1, 10, 1000, 100000):
        x = O.gen_lognormal_f32(n)
        hi, lo = O.sum_dd(x)
        assert hi == math.fsum(float(t) for t in x)
    d = RNG.normal(size=1000) * 1e10
    assert O.sum_dd(d)[0] == math.fsum(d)


def test_segmented_scans_match_groupby():
    keys = np.repeat(np.arange(50, dtype=np.int32), RNG.integers(1, 9, 50))
    vals = RNG.integers(-50, 50, keys.size).astype(np.int32)
    inc = O.segmented_scan(keys, vals, False)
    exc = O.segmented_scan(keys, vals, True, 3)
    i = 0
    for _, grp in itertools.groupby(range(keys.size), key=lambda j: keys[j]):
        idx = list(grp)
        run = np.cumsum(vals[idx])
        assert np.array_equal(inc[idx], run)
        assert np.array_equal(exc[idx], np.r_[0, run[:-1]] + 3)
        i += len(idx)
    x = O.gen_lognormal_f32(keys.size)
    dd = O.segmented_scan_f32_dd(keys, x, False)
    for _, grp in itertools.groupby(range(keys.size), key=lambda j: keys[j]):
        idx = list(grp)
        for t, j in enumerate(idx):
            assert dd[j] == math.fsum(float(x[m]) for m in idx[:t + 1])


# --------------------------------------------------------------- reduction
def test_golden_reduce_and_accumulate(golden):
    n = 0
    opmap = {"plus": O.PLUS, "multiplies": O.MULTIPLIES, "min": O.MIN, "max": O.MAX}
    for c in golden:
        if c["algo"] == "reduce":
            got = O.reduce(_np(c["input"], c["dtype"]), opmap[c["op"]])
        elif c["algo"] == "accumulate":
            x = np.r_[np.array([c["init"]], dtype=c["dtype"]), _np(c["input"], c["dtype"])]
            got = O.reduce(x, opmap[c["op"]])  # left fold from init
        else:
            continue
        assert int(got) == c["expected"], c["id"]
        n += 1
    assert n >= 15


def test_reference_association_orders_closed_forms():
    big = np.float32(2 ** 24)
    # GPU tree (reduce_on_gpu.hpp:186-212): lane l folds x[l + 128 i]; lanes then
    # combine pairwise.  1s in lanes 2 and 3 meet first and survive: 2^24 + 2.
    x = np.zeros(1024, np.float32)
    x[0], x[2], x[3] = big, 1, 1
    assert O.reduce_gpu_order_f32(x) == big + 2
    # the same 1s strided into lane 0 are absorbed one at a time: 2^24.
    x = np.zeros(1024, np.float32)
    x[0], x[128], x[256] = big, 1, 1
    assert O.reduce_gpu_order_f32(x) == big
    # two groups, then the final pass adds the two partials
    x = np.zeros(2048, np.float32)
    x[0], x[1024] = big, 1
    x[1025] = 1  # lane 1 of group 1 -> group partial 2 -> 2^24 + 2
    assert O.reduce_gpu_order_f32(x) == big + 2
    # CPU column (reduce_on_cpu.hpp:66-94): cu chunks, each a left fold
    x = np.zeros(65536, np.float32)
    x[0], x[1], x[32768], x[32769] = big, 1, 1, 1
    assert O.reduce_cpu_order_f32(x, 2) == big + 2
    assert O.reduce_cpu_order_f32(x[:1000], 2) == big  # below threshold: serial
    # accumulate (serial_accumulate.hpp:38-40): left fold from init
    assert O.accumulate_f32(np.ones(2, np.float32), float(big)) == big
    # Hillis-Steele block scan (scan_on_gpu.hpp:84-93): s[2] = (x2+x1) + x0
    x = np.zeros(512, np.float32)
    x[0], x[1], x[2] = big, 1, 1
    inc = O.scan_gpu_order_f32(x, False)
    assert inc[1] == big and inc[2] == big + 2
    # lane 3 = (x3+x2) + (x1+x0) = 1 + 2^24 -> 2^24, and that is the path of
    # the block sum; block 1 adds it: 2^24 + 1 -> 2^24
    assert inc[3] == big
    x[256] = 1
    assert O.scan_gpu_order_f32(x, False)[256] == big
    # with x2 = x3 = 1 the block sum is (x3+x2) + (x1+x0) = 2^24 + 2, and
    # block 1's first element is (2^24 + 2) + 1 -> 2^24 + 4 (ties to even)
    x = np.zeros(512, np.float32)
    x[0], x[2], x[3], x[256] = big, 1, 1, 1
    inc = O.scan_gpu_order_f32(x, False)
    assert inc[3] == big + 2 and inc[256] == big + 4


@pytest.mark.parametrize("cu", [2, 3, 8, 64, 255])
@pytest.mark.parametrize("n", [65535, 65536, 65537, 100003, 196608])
def test_scan_cpu_order_integer_exact(cu, n):
    # int32 plus is exact (wrapping), so every chunk boundary, seed and
    # partial of scan_on_cpu (scan_on_cpu.hpp:75-183) must reproduce the
    # plain prefix sum; an off-by-one in a chunk or seed changes it
    rng = np.random.default_rng(n + cu)
    x = rng.integers(-(1 << 30), 1 << 30, n).astype(np.int32)
    inc = np.cumsum(x.astype(np.int64)).astype(np.int64)
    want_inc = (inc & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    assert np.array_equal(O.scan_cpu_order(x, False, cu), want_inc)
    init = -12345
    exc = np.concatenate([[0], inc[:-1]]) + init
    want_exc = (exc & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    assert np.array_equal(O.scan_cpu_order(x, True, cu, init=init), want_exc)


def test_scan_cpu_order_float_closed_forms():
    big = np.float32(2 ** 24)
    # below max(16384 * sizeof(float), cu) elements: serial_scan, one left
    # fold (numpy's add.accumulate is a sequential float32 left fold)
    rng = np.random.default_rng(5)
    x = rng.normal(size=60000).astype(np.float32)
    assert np.array_equal(O.scan_cpu_order(x, False, 8), np.add.accumulate(x, dtype=np.float32))
    # cu = 2, n = 3 * 65536: block = (n + 2) / 3 = 65536, chunks c0 c1 c2.
    # c0 = [3, 0, ...], c1 = [2^24, 1.5 x 65535], c2 = zeros.  Each 1.5 added
    # at [2^24, 2^25) (ulp 2) rounds up by 2, so c1's own fold is
    # p1 = 2^24 + 131070 = 16908286 and c2's seed is 3 + p1 = 16908289, a tie
    # rounded to even: 16908288.  The serial fold instead gives 3 + 2^24 ->
    # 16777220 (tie to even), then + 131070 = 16908290.
    n = 3 * 65536
    x = np.zeros(n, np.float32)
    x[0] = 3
    x[65536] = big
    x[65537:131072] = 1.5
    inc = O.scan_cpu_order(x, False, 2)
    assert inc[65535] == 3
    assert inc[65536] == 16777220 and inc[131071] == 16777220 + 2 * 65535
    assert np.all(inc[131072:] == 16908288)
    assert np.add.accumulate(x, dtype=np.float32)[131072] == 16908290  # the serial fold differs
    exc = O.scan_cpu_order(x, True, 2, init=0.0)
    assert exc[0] == 0 and exc[1] == 3 and exc[65536] == 3 and exc[65537] == 16777220
    assert exc[131072] == 16908288
    # with one compute unit kernel 2 never runs (scan_on_cpu.hpp:135-137):
    # block = (n + 1) / 2 and the second half keeps what was there
    y = np.ones(131072, np.float32)
    out = O.scan_cpu_order(y, False, 1, fill=-1.0)
    assert np.array_equal(out[:65536], np.arange(1, 65537, dtype=np.float32))
    assert np.all(out[65536:] == -1.0)
    # a work-item whose chunk is empty leaves its partial unwritten (:79):
    # cu = 1000, n = 65536 -> block 66, chunks from 993 on are empty
    with pytest.raises(ValueError):
        O.scan_cpu_order(np.ones(65536, np.float32), False, 1000)


# ------------------------------------------------------------ reduce_by_key
def test_golden_reduce_by_key(golden):
    n = 0
    opmap = {"plus": O.PLUS, "min": O.MIN, "max": O.MAX}
    for c in golden:
        if c["algo"] != "reduce_by_key":
            continue
        k = _np(c["keys"], c["dtype"])
        v = _np(c["values"], c["vdtype"])
        gk, gv = O.reduce_by_key(k, v, opmap[c["op"]])
        assert np.array_equal(gk, _np(c["expected_keys"], c["dtype"])), c["id"]
        if "rtol" in c:
            assert np.allclose(gv, c["expected_values"], rtol=c["rtol"])
        else:
            assert np.array_equal(gv, _np(c["expected_values"], c["vdtype"])), c["id"]
        n += 1
    assert n >= 6


def test_reduce_by_key_matches_groupby_and_writes_last_key():
    keys = np.repeat(RNG.integers(-5, 5, 300).astype(np.int32), RNG.integers(1, 6, 300))
    vals = RNG.integers(-100, 100, keys.size).astype(np.int32)
    gk, gv = O.reduce_by_key(keys, vals)
    ek, ev = [], []
    pos = 0
    for key, grp in itertools.groupby(keys):
        m = len(list(grp))
        ek.append(key)
        ev.append(int(vals[pos:pos + m].sum()))
        pos += m
    assert list(gk) == ek and list(gv) == ev
    # equal_to on floats: -0.0 == +0.0 merges the run and the LAST key is
    # written; NaN never equals, so each NaN is its own run.
    fk = np.array([0.0, -0.0, np.nan, np.nan, 1.0], np.float32)
    gk, gv = O.reduce_by_key(fk, np.ones(5, np.float32))
    assert gk.size == 4 and np.signbit(gk[0]) and np.isnan(gk[1]) and np.isnan(gk[2])
    assert list(gv) == [2.0, 1.0, 1.0, 1.0]
    # double-double run sums are the correctly rounded sums
    fk = np.repeat(np.arange(30, dtype=np.int32), 7)
    fv = O.gen_lognormal_f32(fk.size)
    k2, s2 = O.reduce_by_key_f32_dd(fk, fv)
    assert list(k2) == list(range(30))
    for r in range(30):
        assert s2[r] == math.fsum(float(t) for t in fv[7 * r:7 * r + 7])


# --------------------------------------------------------------- transform
def _round_f32(fr):
    """Correctly round a Fraction to float32 (ties to even), independently of
    any float arithmetic."""
    d = np.float32(float(fr))
    cands = [np.nextafter(d, np.float32(-np.inf)), d, np.nextafter(d, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - fr),
                                     int(np.array(c).view(np.uint32)) & 1))
    return best


def test_saxpy_is_the_fused_multiply_add():
    x, y = O.gen_saxpy_f32(2000)
    got = O.saxpy(2.5, x, y)
    for i in range(2000):
        exact = Fraction(2.5) * Fraction(float(x[i])) + Fraction(float(y[i]))
        assert got[i] == _round_f32(exact)


def test_golden_transform(golden):
    for c in golden:
        if c["algo"] != "transform":
            continue
        op = {"plus": O.PLUS, "multiplies": O.MULTIPLIES}[c["op"]]
        got = O.transform(_np(c["a"], c["dtype"]), _np(c["b"], c["dtype"]), op)
        assert list(got) == c["expected"], c["id"]


# ------------------------------------------------------------- selection
def _pred_of(case):
    p = case["pred"]
    return O.pred(case["dtype"], p["cmp"], p["c"], p["fn"], p["a"], p["negate"])


def test_golden_selection(golden):
    seen = set()
    for c in golden:
        algo = c["algo"]
        if algo not in ("copy_if", "copy_index_if", "unique", "unique_copy", "partition", "remove", "count_if"):
            continue
        seen.add(algo)
        x = np.array(c["input"], dtype=c["dtype"])
        if algo == "copy_if":
            assert O.copy_if(x, _pred_of(c)).tolist() == c["expected"], c["id"]
        elif algo == "copy_index_if":
            assert O.copy_if(x, _pred_of(c), 1).tolist() == c["expected"], c["id"]
        elif algo == "count_if":
            assert O.count_if(x, _pred_of(c)) == c["expected"], c["id"]
        elif algo == "unique_copy":
            assert O.unique_copy(x).tolist() == c["expected"], c["id"]
        elif algo == "unique":
            u = O.unique_copy(x)
            d = x.copy()
            d[:u.size] = u  # unique.hpp: unique_copy from a temporary back into the range
            assert u.size == c["expected_count"], c["id"]
            if "expected" in c:
                assert np.array_equal(d, np.array(c["expected"], c["dtype"])), c["id"]
            if "expected_prefix" in c:
                assert np.array_equal(u, np.array(c["expected_prefix"], c["dtype"])), c["id"]
        elif algo == "partition":
            d, k = O.stable_partition(x, _pred_of(c))
            assert k == c["expected_count"], c["id"]
            assert np.array_equal(d, np.array(c["expected"], c["dtype"])), c["id"]
        elif algo == "remove":
            d = x.copy()
            for s in c["steps"]:
                p = O.pred(c["dtype"], "==", s["value"], negate=True)
                kept = O.copy_if(d[:s["len"]], p)
                d[:kept.size] = kept
                assert kept.size == s["count"] and kept.tolist() == s["expected"], (c["id"], s)
    assert seen == {"copy_if", "copy_index_if", "unique", "unique_copy", "partition", "remove", "count_if"}


def _c_rem(x, a):
    """C's truncating remainder on Python ints."""
    r = abs(x) % abs(a)
    return -r if x < 0 else r


def _wrap(v, bits, signed):
    v &= (1 << bits) - 1
    return v - (1 << bits) if signed and v >> (bits - 1) else v


PRED_CASES = [
    # (dtype, fn, a, cmp, c): chosen so that floor-mod, unsigned-vs-signed,
    # missing integer promotion or a swapped comparison give another answer
    ("int32", "mod", 2, "==", 1), ("int32", "mod", 2, "==", -1), ("int32", "mod", 3, "!=", 0),
    ("int32", "mod", -4, "<", 0), ("int64", "mod", 7, ">=", 3), ("int8", "mod", 4, "==", -1),
    ("uint32", "mod", 6, "==", 5), ("uint64", "mod", 1 << 40, ">", 5),
    ("int8", "mul", 3, ">=", 200), ("uint8", "mul", 2, ">", 255), ("int16", "mul", 3, "<", -40000),
    ("int32", "mul", 3, ">", 0), ("uint32", "id", 0, ">", 2 ** 31), ("uint32", "id", 0, "<=", 7),
    ("int32", "and", 1, "!=", 0), ("uint16", "and", 0x8001, "==", 0x8000), ("int64", "id", 0, "<", -5),
    ("uint64", "id", 0, ">=", 2 ** 63),
]


@pytest.mark.parametrize("case", PRED_CASES, ids=lambda c: "-".join(map(str, c)))
@pytest.mark.parametrize("negate", [False, True])
def test_predicate_integers_follow_c_arithmetic(case, negate):
    dt, fn, a, cmp, c = case
    info = np.iinfo(dt)
    x = RNG.integers(info.min, info.max, size=600, dtype=dt, endpoint=True)
    x[:8] = [info.min, info.max, 0, 1, info.min + 1, info.max - 1, 2, 3] if info.min < 0 else \
        [0, info.max, 1, 2, info.max - 1, 3, 4, 5]
    signed = info.min < 0 or np.dtype(dt).itemsize < 4
    bits = max(32, np.dtype(dt).itemsize * 8)
    want = []
    for v in x.tolist():
        y = v
        if fn == "mul":
            y = _wrap(v * a, bits, signed)
        elif fn == "mod":
            y = _c_rem(v, a)
        elif fn == "and":
            y = _wrap(v & _wrap(a, bits, signed), bits, signed)
        cc = _wrap(c, bits, signed)
        r = {"<": y < cc, "<=": y <= cc, ">": y > cc, ">=": y >= cc, "==": y == cc, "!=": y != cc}[cmp]
        want.append(r != negate)
    got = O.pred_eval(x, O.pred(dt, cmp, c, fn, a, negate))
    assert got.tolist() == want


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_predicate_floats_follow_ieee(dt):
    x = RNG.normal(0, 10, 500).astype(dt)
    x[:6] = [np.nan, -np.nan, 0.0, -0.0, np.inf, -np.inf]
    x[6:10] = np.copysign(np.nan, -1.0), 2.0, 2.5, -2.5
    for cmp in ("<", "<=", ">", ">=", "==", "!="):
        for fn, a in (("id", 0), ("mul", 0.5)):
            for neg in (False, True):
                y = x * np.array(a, dt) if fn == "mul" else x
                with np.errstate(invalid="ignore"):
                    r = {"<": y < 2.5, "<=": y <= 2.5, ">": y > 2.5, ">=": y >= 2.5, "==": y == 2.5,
                         "!=": y != 2.5}[cmp]
                got = O.pred_eval(x, O.pred(dt, cmp, 2.5, fn, a, neg))
                assert np.array_equal(got, r != neg), (cmp, fn, neg)
    sb = O.pred_eval(x, O.pred(dt, "!=", 0, "signbit"))
    assert np.array_equal(sb, np.signbit(x))  # -0.0 and -NaN included


def _bits_equal(a, b):
    """Equal dtype, shape and bits (NaN payloads and signed zeros included)."""
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("dt", ["int8", "uint16", "int32", "uint64", "float32", "float64"])
def test_selection_matches_numpy(dt):
    for n in (0, 1, 31, 32, 1000):
        if np.dtype(dt).kind == "f":
            x = RNG.integers(-3, 4, n).astype(dt)
            if n > 10:
                x[3], x[4], x[7], x[8] = np.nan, np.nan, 0.0, -0.0
            p = O.pred(dt, ">", 0.5)
            mask = x > 0.5
        else:
            x = RNG.integers(0, 5, n).astype(dt)
            p = O.pred(dt, "==", 1, "mod", 2)
            mask = x % 2 == 1
        assert _bits_equal(O.copy_if(x, p), x[mask])
        assert np.array_equal(O.copy_if(x, p, 1), np.flatnonzero(mask).astype(np.uint32))
        assert np.array_equal(O.copy_if(x, p, 2), np.flatnonzero(mask).astype(np.uint64))
        assert O.count_if(x, p) == int(mask.sum())
        t, f = O.partition_copy(x, p)
        assert _bits_equal(t, x[mask]) and _bits_equal(f, x[~mask])
        d, k = O.stable_partition(x, p)
        assert k == int(mask.sum()) and _bits_equal(d, np.concatenate([x[mask], x[~mask]]))
        head = np.ones(n, bool)
        head[1:] = x[1:] != x[:-1]  # != on floats: NaN heads its own run, -0 == +0
        assert _bits_equal(O.unique_copy(x), x[head]), n


# ------------------------------------------------- transform_reduce
def _seq_of(spec, dtype):
    if isinstance(spec, dict):
        return ("count", spec["count"])
    return np.array(spec, dtype)


def test_golden_transform_reduce(golden):
    seen = set()
    for c in golden:
        if c["algo"] not in ("transform_reduce", "inner_product"):
            continue
        seen.add(c["algo"])
        dt = np.dtype(c["dtype"])
        a = _seq_of(c["a"], dt)
        n = c["a"]["n"] if isinstance(c["a"], dict) else len(c["a"])
        if c["algo"] == "inner_product":
            got = O.map_reduce(dt, a, _seq_of(c["b"], dt), n, O.mapping(dt), O.PLUS, c["init"])
            assert got == c["expected"], c["id"]
        else:
            m = O.mapping(dt, c["map"]["fn"], c=c["map"]["c"])
            if dt.kind == "f":
                hi, lo = O.map_sum_dd(dt, a, None, n, m)
                assert abs(hi + lo - c["expected"]) <= c["rtol"] * abs(c["expected"]), c["id"]
            else:
                assert O.map_reduce(dt, a, None, n, m, O.PLUS) == c["expected"], c["id"]
    assert seen == {"transform_reduce", "inner_product"}


def _wrapv(v, dt):
    bits = np.dtype(dt).itemsize * 8
    v &= (1 << bits) - 1
    return v - (1 << bits) if np.dtype(dt).kind == "i" and v >> (bits - 1) else v


@pytest.mark.parametrize("dt", ["int8", "uint16", "int32", "uint32", "int64", "uint64"])
def test_map_reduce_integers_wrap_like_python(dt):
    info = np.iinfo(dt)
    a = RNG.integers(info.min, info.max, 300, endpoint=True, dtype=dt)
    b = RNG.integers(info.min, info.max, 300, endpoint=True, dtype=dt)
    a[:3] = [info.min, info.max, 0]
    cases = [("multiplies", "id", 0), ("minus", "abs", 0), ("plus", "square", 0), (None, "abs", 0),
             (None, "square", 7), (None, "id", -3 if info.min < 0 else 3)]
    for op, fn, cc in cases:
        m = O.mapping(dt, fn, op or "plus", cc)
        for rop, fold in ((O.PLUS, lambda s, t: s + t), (O.MULTIPLIES, lambda s, t: s * t),
                          (O.MIN, min), (O.MAX, max)):
            ys = []
            for x, y in zip(a.tolist(), b.tolist()):
                if op is None:
                    t = _wrapv(x - _wrapv(cc, dt), dt)
                elif op == "multiplies":
                    t = _wrapv(x * y, dt)
                elif op == "minus":
                    t = _wrapv(x - y, dt)
                else:
                    t = _wrapv(x + y, dt)
                if fn == "abs":
                    t = _wrapv(-t, dt) if t < 0 else t  # abs(INT_MIN) keeps INT_MIN
                elif fn == "square":
                    t = _wrapv(t * t, dt)
                ys.append(t)
            want = ys[0]
            for t in ys[1:]:
                want = _wrapv(fold(want, t), dt)
            got = O.map_reduce(dt, a, b if op else None, a.size, m, rop)
            assert int(got) == want, (op, fn, cc, rop)


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_map_sum_dd_is_the_exact_sum_of_exact_products(dt):
    a = (RNG.normal(size=200) * 10.0 ** RNG.integers(-3, 4, 200)).astype(dt)
    b = (RNG.normal(size=200) * 10.0 ** RNG.integers(-3, 4, 200)).astype(dt)
    f = Fraction
    exact = sum(f(float(x)) * f(float(y)) for x, y in zip(a, b))
    hi, lo = O.map_sum_dd(dt, a, b, a.size, O.mapping(dt))
    assert f(hi) + f(lo) - exact <= abs(exact) * f(2) ** -100 + f(2) ** -1070
    assert abs(float(f(hi) + f(lo)) - float(exact)) <= abs(float(exact)) * 2 ** -52
    # squares of (x - c), the difference rounded in the element type first
    c = np.array(0.1, dt)
    d = (a - c).astype(dt)
    exact = sum(f(float(t)) * f(float(t)) for t in d)
    hi, lo = O.map_sum_dd(dt, a, None, a.size, O.mapping(dt, "square", c=c))
    assert abs(float(f(hi) + f(lo)) - float(exact)) <= abs(float(exact)) * 2 ** -52
    # counting sequences: sum of i*i for i < 1000
    hi, lo = O.map_sum_dd(dt, ("count", 0), ("count", 0), 1000, O.mapping(dt))
    assert hi + lo == sum(i * i for i in range(1000))


# ------------------------------------------- experimental::sort_by_transform
def sort_by_transform_oracle(records, get, fn, descending):
    """keys = f(component) then the sort_by_key dispatch (insertion sort below
    32, else stable radix order) with the records as values
    (experimental/sort_by_transform.hpp)."""
    rec = np.ascontiguousarray(records)
    keys = (rec if get is None else rec[:, get]).copy()
    if fn == "abs":
        keys = np.abs(keys)  # integer abs wraps (INT_MIN stays); float abs clears the sign bit
    vals = rec.view(np.dtype("V%d" % (rec.itemsize * (1 if rec.ndim == 1 else rec.shape[1])))).reshape(-1).copy()
    _, v = O.sort_dispatch(O.SORT_BY_KEY, keys, vals, descending)
    return np.frombuffer(v.tobytes(), rec.dtype).reshape(rec.shape)


def test_golden_sort_by_transform(golden):
    seen = 0
    for c in golden:
        if c["algo"] != "sort_by_transform":
            continue
        seen += 1
        got = sort_by_transform_oracle(np.array(c["records"], c["dtype"]), c["get"], c["fn"], c["descending"])
        assert got.tolist() == c["expected"], c["id"]
    assert seen == 2


def test_sort_by_transform_oracle_is_a_stable_sort_by_the_key():
    rec = RNG.integers(-50, 50, (500, 2)).astype(np.int32)
    got = sort_by_transform_oracle(rec, 0, "id", False)
    want = rec[np.argsort(rec[:, 0], kind="stable")]
    assert np.array_equal(got, want)
    got = sort_by_transform_oracle(rec[:, 1].copy(), None, "abs", True)
    k = np.abs(rec[:, 1])
    assert np.array_equal(got, rec[:, 1][np.argsort(-k, kind="stable")])


# ------------------------------------------------------------ comparator sorts
def compare_case_arrays(c):
    """A golden compare_sort case as (records array, record_bytes, values array or None)."""
    dt = np.dtype(c["dtype"])
    rec = np.array(c["records"], dt)
    rb = dt.itemsize * (rec.shape[1] if rec.ndim == 2 else 1)
    vals = np.array(c["values"], c["values_dtype"]) if "values" in c else None
    return rec.reshape(-1), rb, vals


def compare_case_cmp(c):
    m = c["cmp"]
    sz = np.dtype(c["dtype"]).itemsize
    return dict(form=m["form"], lhs=m["lhs"] * sz, rhs=m["rhs"] * sz, second=m.get("second", 0) * sz,
                count=m.get("count", 0), greater=m["greater"])


def check_compare_case(c, out, out_vals):
    """The reference test's own checks on a sorted output (flat arrays)."""
    rec, rb, vals = compare_case_arrays(c)
    dt = np.dtype(c["dtype"])
    comps = rb // dt.itemsize
    got = out.reshape(-1, comps) if comps > 1 else out
    if "expected" in c:
        assert got.tolist() == np.array(c["expected"], dt).tolist(), c["id"]
    for pos, x in c.get("expected_x", {}).items():
        assert got[int(pos)][0] == x, (c["id"], pos)
    if "expected_prefix" in c:
        assert got[:len(c["expected_prefix"])].tolist() == c["expected_prefix"], c["id"]
        assert out_vals[:len(c["expected_values_prefix"])].tolist() == c["expected_values_prefix"], c["id"]
    if c.get("sorted"):
        k = compare_case_cmp(c)
        less = O.comparator_less_py(dt, k["form"], k["lhs"], k["rhs"], k["second"], k["count"], k["greater"])
        rows = O._as_rows(out, rb)
        assert not any(less(rows[i + 1], rows[i]) for i in range(rows.shape[0] - 1)), c["id"]
    if c.get("values_sorted"):
        assert np.all(out_vals[1:] >= out_vals[:-1]), c["id"]
    # a permutation of the input, records and values moved together
    pairs = sorted(zip(map(bytes, O._as_rows(rec, rb)), vals.tolist() if vals is not None else [0] * len(out)))
    got_pairs = sorted(zip(map(bytes, O._as_rows(out, rb)),
                           out_vals.tolist() if out_vals is not None else [0] * len(out)))
    assert pairs == got_pairs, c["id"]


def test_golden_compare_sort(golden):
    seen = 0
    for c in golden:
        if c["algo"] != "compare_sort":
            continue
        seen += 1
        rec, rb, vals = compare_case_arrays(c)
        out, out_vals = O.comparator_sort(rec, rb, c["dtype"], values=vals, **compare_case_cmp(c))
        check_compare_case(c, out, out_vals)
    assert seen == 13


CMP_DTYPES = ["int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64", "float32", "float64"]


def _cmp_records(dt, n, comps, rng):
    dt = np.dtype(dt)
    if dt.kind == "f":
        x = rng.integers(-6, 6, (n, comps)).astype(dt) * dt.type(0.5)  # ties, -0 and +0
        x[rng.random((n, comps)) < 0.1] *= dt.type(-1)
        return x
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, (n, comps), endpoint=True, dtype=dt)
    small = rng.integers(max(info.min, -3), 4, (n, comps)).astype(dt)  # ties
    pick = rng.random((n, comps)) < 0.5
    x[pick] = small[pick]
    x[rng.random((n, comps)) < 0.05] = info.min
    x[rng.random((n, comps)) < 0.05] = info.max
    return x


@pytest.mark.parametrize("dt", CMP_DTYPES)
def test_comparator_keys_match_the_insertion_sort(dt):
    # the vectorised key path against the reference's insertion sort driven by
    # the pairwise comparator in exact Python arithmetic (wrap-around, abs of
    # INT_MIN, -0 == +0, promotion of narrow integers)
    rng = np.random.default_rng(99)
    forms = ["field", "abs", "abs_sum2"] + (["length"] if np.dtype(dt).kind == "f" else [])
    sz = np.dtype(dt).itemsize
    for comps in (1, 2, 4):
        for form in forms:
            if form == "abs_sum2" and comps == 1:
                continue
            for greater in (False, True):
                rec = _cmp_records(dt, 97, comps, rng)
                rows = O._as_rows(rec, comps * sz)
                kw = dict(form=form, lhs=sz * (comps - 1 if form == "field" else 0), second=sz * (comps - 1),
                          count=comps, greater=greater)
                vals = np.arange(97, dtype=np.int32)
                out, ov = O.comparator_sort(rec, comps * sz, dt, values=vals, **kw)
                less = O.comparator_less_py(dt, kw["form"], kw["lhs"], kw["lhs"], kw["second"], kw["count"],
                                            greater)
                want, wv = O.insertion_sort_rows(rows, less, list(vals))
                assert bytes(O._as_rows(out, comps * sz)) == bytes(want.astype(np.uint8)), (comps, form, greater)
                assert ov.tolist() == wv, (comps, form, greater)


def test_comparator_oracle_catches_a_wrapped_abs(monkeypatch):
    # a planted mistake: abs(INT_MIN) kept as INT_MIN (a wrapping abs) sorts
    # INT_MIN first instead of last, which the insertion-sort check sees
    rec = np.array([3, -2147483648, -5, 7], np.int32)
    good, _ = O.comparator_sort(rec, 4, np.int32, form="abs")
    assert good.tolist() == [3, -5, 7, -2147483648]
    monkeypatch.setattr(O, "_magnitude_u64", lambda s: np.abs(s).astype(np.int64))
    bad, _ = O.comparator_sort(rec, 4, np.int32, form="abs")
    less = O.comparator_less_py(np.int32, "abs", 0)
    want, _ = O.insertion_sort_rows(O._as_rows(rec, 4), less)
    assert bytes(O._as_rows(bad, 4)) != bytes(want.astype(np.uint8))


def test_comparator_oracle_asymmetric_pairwise():
    # a.s0 < b.s3 on long8_ records whose components are all equal is the
    # ordinary order of the components (test/test_merge_sort_gpu.cpp:299-330)
    rec = np.repeat(np.array([5, -1, 9, 5, 0, 2], np.int64)[:, None], 8, axis=1)
    out, _ = O.comparator_sort(rec, 64, np.int64, form="field", lhs=0, rhs=24)
    assert out.reshape(-1, 8)[:, 0].tolist() == [-1, 0, 2, 5, 5, 9]
