"""Parity of the HIP path (through the C-ABI) with the oracle and with the
reference's own golden vectors.

Bar: bit-exact for sorts, integer scans/reductions and reduce_by_key keys;
float results within the stated bound of the double-double oracle (the
reference's float association is not reproduced: the device sums in FP64 or
compensated pairs, so its error bound is tighter than a sequential float sum's)."""
import math

import numpy as np
import pytest

import compute_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(20261202)
OPS = {"plus": ca.PLUS, "multiplies": ca.MULTIPLIES, "min": ca.MIN, "max": ca.MAX}
OOPS = {"plus": O.PLUS, "multiplies": O.MULTIPLIES, "min": O.MIN, "max": O.MAX}


def dev(q, a):
    return ca.DeviceArray.from_host(np.ascontiguousarray(a), q)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.itemsize])


def _np(vals, dt):
    if dt.startswith("float"):
        return np.array(vals, dtype=dt)
    return np.array(vals, dtype=np.int64).astype(dt)


def random_keys(dt, n, rng=RNG):
    dt = np.dtype(dt)
    if dt.kind == "f":
        k = (rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, n)).astype(dt)
        if n > 8:
            sel = rng.integers(0, n, max(1, n // 50))
            k[sel] = rng.choice(np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1.0], dt),
                                sel.size)
            k[rng.integers(0, n, max(1, n // 20))] = k[rng.integers(0, n, max(1, n // 20))]
        return k
    info = np.iinfo(dt)
    k = rng.integers(info.min, info.max, n, endpoint=True, dtype=dt)
    if n > 8:  # duplicates and extremes
        k[rng.integers(0, n, n // 10)] = k[rng.integers(0, n, n // 10)]
        k[rng.integers(0, n, 3)] = info.min
        k[rng.integers(0, n, 3)] = info.max
    return k


# =========================================================== golden vectors
def test_golden_sorts(queue, golden):
    for c in golden:
        if c["algo"] not in ("radix_sort", "sort", "radix_sort_range"):
            continue
        x = _np(c["input"], c["dtype"])
        d = dev(queue, x)
        desc = c.get("descending", False)
        if c["algo"] == "radix_sort_range":
            f, l = c["first"], c["last"]
            ca.sort(queue, d, api=ca.RADIX_SORT, n=l - f, key_ptr=d.at(f), descending=desc)
        else:
            api = ca.RADIX_SORT if c["algo"] == "radix_sort" else ca.SORT
            ca.sort(queue, d, api=api, descending=desc)
        got = d.to_host(queue)
        exp = _np(c["expected"], c["dtype"])
        assert np.array_equal(got, exp), c["id"]
        # and bit-for-bit the oracle's order (signed zeros included)
        if c["algo"] == "radix_sort":
            assert np.array_equal(bits(got), bits(O.radix_order_sort(x, descending=desc))), c["id"]
        elif c["algo"] == "sort":
            assert np.array_equal(bits(got), bits(O.sort_dispatch(O.SORT, x, descending=desc))), c["id"]


def test_golden_sort_by_key(queue, golden):
    for c in golden:
        if c["algo"] not in ("radix_sort_by_key", "sort_by_key", "stable_sort_by_key_prefix"):
            continue
        k = _np(c["keys"], c["dtype"])
        v = _np(c["values"], c["vdtype"])
        dk, dv = dev(queue, k), dev(queue, v)
        api = {"radix_sort_by_key": ca.RADIX_SORT, "sort_by_key": ca.SORT_BY_KEY,
               "stable_sort_by_key_prefix": ca.STABLE_SORT_BY_KEY}[c["algo"]]
        ca.sort(queue, dk, dv, api=api, descending=c.get("descending", False))
        gk, gv = dk.to_host(queue), dv.to_host(queue)
        if c["algo"] == "stable_sort_by_key_prefix":
            m = len(c["expected_keys_prefix"])
            assert list(gk[:m]) == c["expected_keys_prefix"], c["id"]
            assert list(gv[:m]) == c["expected_values_prefix"], c["id"]
        else:
            assert np.array_equal(gk, _np(c["expected_keys"], c["dtype"])), c["id"]
            assert np.array_equal(gv, _np(c["expected_values"], c["vdtype"])), c["id"]


def test_golden_scans(queue, golden):
    for c in golden:
        if c["algo"] not in ("inclusive_scan", "exclusive_scan"):
            continue
        x = _np(c["input"], c["dtype"])
        excl = c["algo"] == "exclusive_scan"
        for inplace in (False, True):
            d = dev(queue, x)
            out = d if inplace else ca.DeviceArray(x.size, x.dtype, queue)
            ca.scan(queue, d, out, excl, init=c.get("init") if excl else None,
                    op=OPS[c.get("op", "plus")])
            assert np.array_equal(out.to_host(queue), _np(c["expected"], c["dtype"])), c["id"]


def test_golden_reduce_and_accumulate(queue, golden):
    for c in golden:
        if c["algo"] not in ("reduce", "accumulate"):
            continue
        x = _np(c["input"], c["dtype"])
        out = ca.DeviceArray(1, x.dtype, queue)
        init = c.get("init") if c["algo"] == "accumulate" else None
        ca.reduce(queue, dev(queue, x), out, op=OPS[c["op"]], init=init)
        assert int(out.to_host(queue)[0]) == c["expected"], c["id"]


def test_golden_reduce_by_key(queue, golden):
    for c in golden:
        if c["algo"] != "reduce_by_key":
            continue
        k = _np(c["keys"], c["dtype"])
        v = _np(c["values"], c["vdtype"])
        ko = ca.DeviceArray(k.size, k.dtype, queue)
        vo = ca.DeviceArray(k.size, v.dtype, queue)
        r = ca.reduce_by_key(queue, dev(queue, k), dev(queue, v), ko, vo, op=OPS[c["op"]])
        assert r == len(c["expected_keys"]), c["id"]
        assert np.array_equal(ko.to_host(queue)[:r], _np(c["expected_keys"], c["dtype"])), c["id"]
        gv = vo.to_host(queue)[:r]
        if "rtol" in c:
            assert np.allclose(gv, c["expected_values"], rtol=c["rtol"]), c["id"]
        else:
            assert np.array_equal(gv, _np(c["expected_values"], c["vdtype"])), c["id"]


def test_golden_transform(queue, golden):
    for c in golden:
        if c["algo"] != "transform":
            continue
        a, b = _np(c["a"], c["dtype"]), _np(c["b"], c["dtype"])
        out = ca.DeviceArray(a.size, a.dtype, queue)
        ca.transform(queue, dev(queue, a), dev(queue, b), out, OPS[c["op"]])
        assert list(out.to_host(queue)) == c["expected"], c["id"]


# ================================================================== sorts
ALL_KEYS = ["uint8", "int8", "uint16", "int16", "uint32", "int32", "uint64", "int64", "float32",
            "float64"]


@pytest.mark.parametrize("dt", ALL_KEYS)
@pytest.mark.parametrize("n", [2, 17, 32, 33, 300, 4095, 4096, 4097, 12289, 70001])
def test_sort_keys_bit_exact(queue, dt, n):
    k = random_keys(dt, n)
    for api in (ca.SORT, ca.STABLE_SORT):
        for desc in (False, True):
            d = dev(queue, k)
            ca.sort(queue, d, api=api, descending=desc)
            exp = O.sort_dispatch(api, k, descending=desc)
            assert np.array_equal(bits(d.to_host(queue)), bits(exp)), (api, desc)


@pytest.mark.parametrize("dt", ["uint32", "int32", "float32", "uint64"])
@pytest.mark.parametrize("pattern", ["one value", "three values", "top byte fixed", "low byte fixed", "two hot digits"])
def test_sort_skewed_digits(queue, dt, pattern):
    """Skewed digits: few distinct keys, one byte fixed (only that pass
    skewed), two hot digits (same-address LDS atomics in counting and
    ranking); keys only and with values, both directions, checked against
    the oracle's stable radix order."""
    n = 3000017
    u = random_keys("uint64", n).view(np.uint64)
    if pattern == "one value":
        k = np.full(n, 0x1234, np.uint64)
    elif pattern == "three values":
        k = np.array([7, 0xABCD, 0x5555_0000], np.uint64)[RNG.integers(0, 3, n)]
    elif pattern == "top byte fixed":
        k = (u & np.uint64(0x00FF_FFFF)) | np.uint64(0x5A00_0000)
    elif pattern == "low byte fixed":
        k = (u & ~np.uint64(0xFF)) | np.uint64(0x3C)
    else:
        k = np.where(RNG.random(n) < 0.4, np.uint64(0x11), np.uint64(0xEE)) | (u & ~np.uint64(0xFF)) * \
            (RNG.random(n) < 0.5).astype(np.uint64)
    dtn = np.dtype(dt)
    if dtn.kind == "f":
        k = (k & np.uint64(0xFFFF_FFFF)).astype(np.uint32).view(np.float32)
        k = np.where(np.isnan(k), np.float32(1.5), k).astype(np.float32)
    else:
        k = (k & np.uint64((1 << (8 * dtn.itemsize)) - 1)).astype(dtn.str.replace("i", "u")).view(dtn)
    v = np.arange(n, dtype=np.uint32)
    for desc in (False, True):
        d = dev(queue, k)
        ca.sort(queue, d, api=ca.SORT, descending=desc)
        exp = O.sort_dispatch(ca.SORT, k, descending=desc)
        assert np.array_equal(bits(d.to_host(queue)), bits(exp)), (pattern, desc)
        dk, dv = dev(queue, k), dev(queue, v)
        ca.sort(queue, dk, dv, api=ca.SORT_BY_KEY, descending=desc)
        ek, ev = O.sort_dispatch(ca.SORT_BY_KEY, k, v, descending=desc)
        assert np.array_equal(bits(dk.to_host(queue)), bits(ek)), (pattern, desc)
        assert np.array_equal(dv.to_host(queue), ev), (pattern, desc)


@pytest.mark.parametrize("dt", ["uint32", "int32", "float32", "float64", "uint64", "int8", "uint16"])
@pytest.mark.parametrize("vdt", ["uint8", "uint16", "uint32", "uint64"])
@pytest.mark.parametrize("n", [5, 31, 32, 1000, 5000, 40000])
def test_sort_by_key_stable_bit_exact(queue, dt, vdt, n):
    k = random_keys(dt, n)
    if n > 10:
        k[RNG.integers(0, n, n // 3)] = k[0]  # many ties
    v = np.arange(n).astype(vdt)
    for api in (ca.SORT_BY_KEY, ca.STABLE_SORT_BY_KEY):
        for desc in (False, True):
            dk, dv = dev(queue, k), dev(queue, v)
            ca.sort(queue, dk, dv, api=api, descending=desc)
            ek, ev = O.sort_dispatch(api, k, v, descending=desc)
            assert np.array_equal(bits(dk.to_host(queue)), bits(ek)), (api, desc)
            assert np.array_equal(dv.to_host(queue), ev), (api, desc)


def test_sort_by_key_16_byte_values(queue):
    n = 9000
    k = random_keys("uint32", n)
    v = RNG.integers(0, 2 ** 32, (n, 4), dtype=np.uint32)
    dk = dev(queue, k)
    dv = ca.DeviceArray.from_host(v.reshape(-1), queue)
    ca.check(ca.lib().bcmi_sort(queue.h, ca.SORT_BY_KEY, ca.U32, dk.ptr, dv.ptr, 16, n, 0))
    ek, order = O.radix_order_sort(k, np.arange(n, dtype=np.uint32))
    assert np.array_equal(dk.to_host(queue), ek)
    assert np.array_equal(dv.to_host(queue).reshape(n, 4), v[order])


def test_sort_unaligned_subrange(queue):
    k = random_keys("uint32", 100003)
    d = dev(queue, k)
    ca.sort(queue, d, api=ca.SORT, n=100003 - 7 - 3, key_ptr=d.at(7))
    got = d.to_host(queue)
    assert np.array_equal(got[:7], k[:7]) and np.array_equal(got[-3:], k[-3:])
    assert np.array_equal(got[7:-3], np.sort(k[7:-3]))


def test_sort_plain_uint32_config_bit_identical(queue):
    """sort of 2^20 uint32 from mt19937(0xB0057001): bit-identical to the
    reference (the sorted multiset is unique)."""
    k = O.gen_sort_u32(1 << 20)
    d = dev(queue, k)
    ca.sort(queue, d)
    assert np.array_equal(d.to_host(queue), np.sort(k))


def test_sort_by_key_float_config_small(queue):
    """The sort_by_key workload (float keys with signed zeros and duplicates,
    payload = index) at 2^20: keys and payload bit-identical to the oracle."""
    k, v = O.gen_sort_by_key_f32(1 << 20)
    dk, dv = dev(queue, k), dev(queue, v)
    ca.sort(queue, dk, dv, api=ca.SORT_BY_KEY)
    ek, ev = O.radix_order_sort(k, v)
    assert np.array_equal(bits(dk.to_host(queue)), bits(ek))
    assert np.array_equal(dv.to_host(queue), ev)


def test_sort_full_size_properties(queue):
    """2^27 uint32 (the metric's size): sorted, and the same multiset
    (checksums of the raw keys are order-independent)."""
    n = 1 << 27
    k = O.gen_sort_u32(n)
    d = dev(queue, k)
    ca.sort(queue, d)
    got = d.to_host(queue)
    assert np.all(got[1:] >= got[:-1])
    assert int(got.astype(np.uint64).sum()) == int(k.astype(np.uint64).sum())
    assert int(np.bitwise_xor.reduce(got)) == int(np.bitwise_xor.reduce(k))
    h = lambda a: int(((a.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(17)).sum())
    assert h(got) == h(k)


# ================================================================== scans
@pytest.mark.parametrize("dt", ["int32", "uint32", "int64", "uint64", "int8", "uint16", "float64"])
@pytest.mark.parametrize("n", [1, 5, 1023, 4096, 4097, 65536 + 17, (1 << 20) + 3])
def test_scan_exact(queue, dt, n):
    if np.dtype(dt).kind == "f":
        x = RNG.integers(-1000, 1000, n).astype(dt)  # exact in double
    else:
        x = RNG.integers(-100, 100, n).astype(dt)
    d = dev(queue, x)
    out = ca.DeviceArray(n, x.dtype, queue)
    for excl in (False, True):
        init = 5 if excl else None
        ca.scan(queue, d, out, excl, init=init, precision=ca.NAIVE)
        exp = O.scan(x, excl, 5 if excl else 0)
        assert np.array_equal(out.to_host(queue), exp), excl


@pytest.mark.parametrize("op", ["multiplies", "min", "max"])
@pytest.mark.parametrize("dt", ["int32", "uint64", "int16"])
def test_scan_ops(queue, op, dt):
    n = 50001
    x = RNG.integers(-3, 4, n).astype(dt) if op == "multiplies" else random_keys(dt, n)
    d = dev(queue, x)
    out = ca.DeviceArray(n, x.dtype, queue)
    for excl in (False, True):
        init = {"multiplies": 1, "min": 7, "max": 7}[op] if excl else None
        ca.scan(queue, d, out, excl, init=init, op=OPS[op])
        exp = O.scan(x, excl, init if excl else 0, OOPS[op])
        assert np.array_equal(out.to_host(queue), exp), excl


def test_scan_in_place_and_unaligned(queue):
    n = 300007
    x = RNG.integers(-100, 100, n).astype(np.int32)
    d = dev(queue, x)
    ca.scan(queue, d, d, True, init=3)
    assert np.array_equal(d.to_host(queue), O.scan(x, True, 3))
    d = dev(queue, x)
    out = ca.DeviceArray(n, np.int32, queue)
    # input at +1 element, output at +2: neither 16-byte aligned
    m = n - 2
    ca.check(ca.lib().bcmi_scan(queue.h, ca.I32, ca.PLUS, d.at(1), out.at(2), m, None, 0, 0))
    assert np.array_equal(out.to_host(queue)[2:], O.scan(x[1:1 + m], False))


def test_scan_full_size_int32(queue):
    """exclusive_scan of 2^28 int32 in [-100, 100]: bit-exact."""
    n = 1 << 28
    x = O.gen_scan_i32(n)
    d = dev(queue, x)
    out = ca.DeviceArray(n, np.int32, queue)
    ca.scan(queue, d, out, True, init=0)
    got = out.to_host(queue)
    exp = np.empty(n, np.int32)
    exp[0] = 0
    np.cumsum(x[:-1], dtype=np.int32, out=exp[1:])
    assert np.array_equal(got, exp)


def _scan_err_bound(x, dd):
    # compensated within a tile + FP64 carry: |err| <= ulp(result) plus the
    # FP64 carry's rounding, 2^-50 * sum|x| of the prefix
    return np.spacing(np.abs(dd).astype(np.float32)).astype(np.float64) + \
        2.0 ** -50 * np.cumsum(np.abs(x.astype(np.float64)))


@pytest.mark.parametrize("n", [100, 4096 * 3 + 5, 1 << 20])
def test_scan_float_compensated(queue, n):
    x = O.gen_lognormal_f32(n)
    d = dev(queue, x)
    out = ca.DeviceArray(n, np.float32, queue)
    ca.scan(queue, d, out, False)
    dd = O.scan_f32_dd(x, False)
    err = np.abs(out.to_host(queue).astype(np.float64) - dd)
    assert np.all(err <= _scan_err_bound(x, dd))
    ca.scan(queue, d, out, True, init=0.0)
    dd = O.scan_f32_dd(x, True)
    err = np.abs(out.to_host(queue).astype(np.float64) - dd)
    assert np.all(err <= _scan_err_bound(x, dd))


def test_scan_float_compensated_edge_forms(queue):
    """The compensated float scan on partial tiles, exclusive with an init,
    in place, and with unaligned input and output (scalar loads)."""
    L = ca.lib()
    for n in (100, 4096 * 3 + 5, 8192 * 4, (1 << 20) + 17):
        x = O.gen_lognormal_f32(n)
        d = dev(queue, x)
        out = ca.DeviceArray(n, np.float32, queue)
        ca.scan(queue, d, out, False)
        dd = O.scan_f32_dd(x, False)
        assert np.all(np.abs(out.to_host(queue).astype(np.float64) - dd) <= _scan_err_bound(x, dd)), n
        # exclusive with init 2.5: the oracle scans [2.5, x0, x1, ...]
        ca.scan(queue, d, out, True, init=np.float32(2.5))
        xi = np.r_[np.float32(2.5), x[:-1]].astype(np.float32)
        dd = O.scan_f32_dd(xi, False)
        assert np.all(np.abs(out.to_host(queue).astype(np.float64) - dd) <= _scan_err_bound(xi, dd)), n
        # in place
        ca.scan(queue, d, d, False)
        dd = O.scan_f32_dd(x, False)
        assert np.all(np.abs(d.to_host(queue).astype(np.float64) - dd) <= _scan_err_bound(x, dd)), n
        # input at +1 element, output at +2
        d = dev(queue, x)
        m = n - 2
        ca.check(L.bcmi_scan(queue.h, ca.F32, ca.PLUS, d.at(1), out.at(2), m, None, 0, 0))
        dd = O.scan_f32_dd(np.ascontiguousarray(x[1:1 + m]), False)
        got = out.to_host(queue)[2:].astype(np.float64)
        assert np.all(np.abs(got - dd) <= _scan_err_bound(x[1:1 + m], dd)), n


# ================================================================= reduce
@pytest.mark.parametrize("dt", ["int32", "uint32", "int64", "uint8", "int16", "uint64"])
@pytest.mark.parametrize("op", ["plus", "multiplies", "min", "max"])
@pytest.mark.parametrize("n", [1, 7, 1000, 1 << 20])
def test_reduce_integers_exact(queue, dt, op, n):
    x = RNG.integers(-3, 4, n).astype(dt) if op == "multiplies" else random_keys(dt, n)
    out = ca.DeviceArray(1, x.dtype, queue)
    ca.reduce(queue, dev(queue, x), out, op=OPS[op])
    assert out.to_host(queue)[0] == O.reduce(x, OOPS[op])
    # accumulate = fold with init
    ini = np.array([2], dtype=dt)
    ca.reduce(queue, dev(queue, x), out, op=OPS[op], init=ini[0])
    assert out.to_host(queue)[0] == O.reduce(np.r_[ini, x], OOPS[op])


def test_reduce_unaligned(queue):
    x = RNG.integers(-100, 100, 100001).astype(np.int32)
    d = dev(queue, x)
    out = ca.DeviceArray(1, np.int32, queue)
    ca.check(ca.lib().bcmi_reduce(queue.h, ca.I32, ca.PLUS, d.at(3), 100001 - 5, None, out.ptr, None, 0))
    assert out.to_host(queue)[0] == x[3:-2].sum()


@pytest.mark.parametrize("n", [1, 33, 4096, 1 << 20, (1 << 22) + 11])
def test_reduce_float_compensated_within_half_ulp_of_fp64(queue, n):
    x = O.gen_lognormal_f32(n)
    out = ca.DeviceArray(1, np.float32, queue)
    f64 = ca.reduce(queue, dev(queue, x), out)
    got = float(out.to_host(queue)[0])
    hi, lo = O.sum_dd(x)
    exact = hi + lo
    ulp = float(np.spacing(np.float32(abs(exact))))
    assert abs(got - exact) <= 0.5 * ulp + 1e-30
    assert abs(float(f64.to_host(queue)[0]) - exact) <= 2.0 ** -45 * float(np.abs(x.astype(np.float64)).sum())


def test_reduce_float_full_size_within_two_ulp(queue):
    """accumulate/reduce(plus) on 2^28 lognormal(0,3) floats with random
    signs: the compensated result is within 2 ulp of the FP64 oracle."""
    n = 1 << 28
    x = O.gen_lognormal_f32(n)
    out = ca.DeviceArray(1, np.float32, queue)
    ca.reduce(queue, dev(queue, x), out)
    got = float(out.to_host(queue)[0])
    hi, lo = O.sum_dd(x)
    ulp = float(np.spacing(np.float32(abs(hi))))
    assert abs(got - (hi + lo)) <= 2 * ulp


def test_reduce_double_and_empty(queue):
    x = RNG.normal(size=100000)
    out = ca.DeviceArray(1, np.float64, queue)
    ca.reduce(queue, dev(queue, x), out)
    assert abs(out.to_host(queue)[0] - math.fsum(x)) <= 4 * np.spacing(abs(math.fsum(x)))
    # accumulate of an empty range returns init
    ca.check(ca.lib().bcmi_reduce(queue.h, ca.F64, ca.PLUS, None, 0,
                                  np.array([2.5]).ctypes.data_as(ca.C.c_void_p), out.ptr, None, 0))
    assert out.to_host(queue)[0] == 2.5


# ========================================================== reduce_by_key
@pytest.mark.parametrize("n", [1, 2, 257, 4096, 4097, 100003, 1 << 20])
def test_reduce_by_key_int_exact(queue, n):
    lens = RNG.geometric(1 / 20, n)
    keys = np.repeat(np.arange(lens.size, dtype=np.int32), lens)[:n]
    keys = (keys * 7 - 1000).astype(np.int32)
    vals = RNG.integers(-1000, 1000, n).astype(np.int32)
    ko = ca.DeviceArray(n, np.int32, queue)
    vo = ca.DeviceArray(n, np.int32, queue)
    for op in ("plus", "min", "max"):
        r = ca.reduce_by_key(queue, dev(queue, keys), dev(queue, vals), ko, vo, op=OPS[op])
        ek, ev = O.reduce_by_key(keys, vals, OOPS[op])
        assert r == ek.size
        assert np.array_equal(ko.to_host(queue)[:r], ek)
        assert np.array_equal(vo.to_host(queue)[:r], ev)


def test_reduce_by_key_float_keys_signed_zero_and_nan(queue):
    k = np.array([1.0, 0.0, -0.0, 0.0, np.nan, np.nan, 2.0, 2.0] * 700, np.float32)
    v = np.ones(k.size, np.int64)
    ko = ca.DeviceArray(k.size, np.float32, queue)
    vo = ca.DeviceArray(k.size, np.int64, queue)
    r = ca.reduce_by_key(queue, dev(queue, k), dev(queue, v), ko, vo)
    ek, ev = O.reduce_by_key(k, v)
    assert r == ek.size
    assert np.array_equal(bits(ko.to_host(queue)[:r]), bits(ek))
    assert np.array_equal(vo.to_host(queue)[:r], ev)


@pytest.mark.parametrize("n", [1000, 1 << 20])
def test_reduce_by_key_float_compensated(queue, n):
    keys, vals = O.gen_runs_i32_f32(n)
    ko = ca.DeviceArray(n, np.int32, queue)
    vo = ca.DeviceArray(n, np.float32, queue)
    r = ca.reduce_by_key(queue, dev(queue, keys), dev(queue, vals), ko, vo)
    ek, es = O.reduce_by_key_f32_dd(keys, vals)
    assert r == ek.size
    assert np.array_equal(ko.to_host(queue)[:r], ek)
    got = vo.to_host(queue)[:r].astype(np.float64)
    # each run: correctly rounded up to the FP64 carry's rounding
    absum = np.add.reduceat(np.abs(vals.astype(np.float64)), np.r_[0, np.flatnonzero(np.diff(keys)) + 1])
    bound = np.spacing(np.abs(es).astype(np.float32)).astype(np.float64) + 2.0 ** -50 * absum
    assert np.all(np.abs(got - es) <= bound)


# ========================================================= segmented scan
@pytest.mark.parametrize("n", [1, 100, 4097, 200003])
def test_segmented_scan_int_exact(queue, n):
    lens = RNG.integers(1, 70, n)
    keys = np.repeat(np.arange(lens.size, dtype=np.int32), lens)[:n]
    vals = RNG.integers(-100, 100, n).astype(np.int32)
    out = ca.DeviceArray(n, np.int32, queue)
    for excl in (False, True):
        ca.segmented_scan(queue, dev(queue, keys), dev(queue, vals), out, excl, init=4 if excl else None)
        assert np.array_equal(out.to_host(queue), O.segmented_scan(keys, vals, excl, 4 if excl else 0))


def test_segmented_scan_float(queue):
    keys, vals = O.gen_runs_i32_f32(300001)
    out = ca.DeviceArray(vals.size, np.float32, queue)
    for excl in (False, True):
        ca.segmented_scan(queue, dev(queue, keys), dev(queue, vals), out, excl, init=0.0 if excl else None)
        dd = O.segmented_scan_f32_dd(keys, vals, excl)
        err = np.abs(out.to_host(queue).astype(np.float64) - dd)
        bound = np.spacing(np.abs(dd).astype(np.float32)).astype(np.float64) + \
            2.0 ** -50 * np.cumsum(np.abs(vals.astype(np.float64)))
        assert np.all(err <= bound)


# ============================================================== transform
@pytest.mark.parametrize("n", [1, 3, 1000, (1 << 20) + 7])
def test_saxpy_bit_exact(queue, n):
    x, y = O.gen_saxpy_f32(n)
    out = ca.DeviceArray(n, np.float32, queue)
    ca.saxpy(queue, 2.5, dev(queue, x), dev(queue, y), out)
    assert np.array_equal(bits(out.to_host(queue)), bits(O.saxpy(2.5, x, y)))


def test_saxpy_unaligned_in_place(queue):
    n = 100001
    x, y = O.gen_saxpy_f32(n)
    dx, dy = dev(queue, x), dev(queue, y)
    m = n - 5
    ca.check(ca.lib().bcmi_saxpy(queue.h, 2.5, dx.at(1), dy.at(1), dy.at(1), m))
    exp = y.copy()
    exp[1:1 + m] = O.saxpy(2.5, x[1:1 + m], y[1:1 + m])
    assert np.array_equal(bits(dy.to_host(queue)), bits(exp))
    ca.check(ca.lib().bcmi_saxpy(queue.h, 2.5, dx.at(1), dx.at(2), dy.ptr, m))  # mutually unaligned
    assert np.array_equal(bits(dy.to_host(queue)[:m]), bits(O.saxpy(2.5, x[1:1 + m], x[2:2 + m])))


@pytest.mark.parametrize("dt", ["int32", "uint8", "int64", "float32", "float64"])
@pytest.mark.parametrize("op", ["plus", "multiplies", "min", "max"])
def test_transform_binary_ops(queue, dt, op):
    n = 12345
    a, b = random_keys(dt, n), random_keys(dt, n)
    if np.dtype(dt).kind == "f":
        a, b = np.nan_to_num(a), np.nan_to_num(b)
    out = ca.DeviceArray(n, a.dtype, queue)
    ca.transform(queue, dev(queue, a), dev(queue, b), out, OPS[op])
    assert np.array_equal(bits(out.to_host(queue)), bits(O.transform(a, b, OOPS[op])))


# ============================================================ hipGraph
def test_graph_replay_matches_eager(queue):
    n = 300000
    k = random_keys("uint32", n)
    x = RNG.integers(-100, 100, n).astype(np.int32)
    dk, dx = dev(queue, k), dev(queue, x)
    out = ca.DeviceArray(n, np.int32, queue)
    queue.reserve(n, 8)
    g = ca.Graph(queue)
    with g:
        ca.sort(queue, dk)
        ca.scan(queue, dx, out, True, init=0)
    exp_k, exp_s = np.sort(k), O.scan(x, True, 0)
    for rep in range(3):
        ca.check(ca.lib().bcmi_memcpy(queue.h, dk.ptr, k.ctypes.data_as(ca.C.c_void_p), k.nbytes, ca.H2D, 1))
        g.launch()
        queue.finish()
        assert np.array_equal(dk.to_host(queue), exp_k), rep
        assert np.array_equal(out.to_host(queue), exp_s), rep
        # eager work between replays
        ca.scan(queue, dx, out, False)
        queue.finish()


def test_graph_replay_after_arena_growth_is_refused():
    """A graph holds its queue's scratch arena and status array by address.
    When a later eager call grows them (the old buffers are freed), a replay
    is refused with an error instead of touching freed memory; a graph
    captured again afterwards replays correctly.  A graph only replays on
    the queue it was captured on."""
    q = ca.Queue(0)
    other = ca.Queue(0)
    try:
        n_small, n_big = 5000, 3 << 20
        k = random_keys("uint32", n_small)
        dk = dev(q, k)
        ca.sort(q, dk)  # sizes the arena for n_small
        q.finish()
        g = ca.Graph(q)
        with g:
            ca.sort(q, dk)
        with pytest.raises(ca.BcmiError):
            g.launch(other)
        big = dev(q, random_keys("uint32", n_big))
        ca.sort(q, big)  # grows the arena and the status array
        q.finish()
        with pytest.raises(ca.BcmiError, match="reallocated"):
            g.launch()
        g2 = ca.Graph(q)
        with g2:
            ca.sort(q, dk)
        for rep in range(2):
            ca.check(ca.lib().bcmi_memcpy(q.h, dk.ptr, k.ctypes.data_as(ca.C.c_void_p), k.nbytes, ca.H2D, 1))
            g2.launch()
            q.finish()
            assert np.array_equal(dk.to_host(q), np.sort(k)), rep
        del g, g2
    finally:
        other.close()
        q.close()


# ================================================ tunable radix variants
def _tune_get(key, param):
    v = ca.C.c_uint()
    ca.check(ca.lib().bcmi_tune_get(key.encode(), param.encode(), 0xFFFFFFFF, ca.C.byref(v)))
    return v.value


@pytest.mark.parametrize("threads,ipt,split", [(256, 16, 1), (256, 32, 1), (256, 64, 1), (512, 16, 1), (512, 24, 1),
                                               (512, 32, 1), (1024, 16, 1), (256, 64, 2), (512, 24, 2),
                                               (512, 32, 2), (512, 24, 3)])
def test_sort_tunable_tile_variants(queue, threads, ipt, split):
    """Every tile shape the tuner may select sorts exactly, keys only and
    with values, including partial last tiles, several histogram shapes,
    XCD-grouped tile orders (full groups and a trailing partial group) and
    descending order."""
    params = ("threads", "ipt", "fit_size", "xcd_run", "split")
    defaults = {"threads": 512, "ipt": 32, "fit_size": 1, "xcd_run": 0, "split": 1}
    saved = {k: [_tune_get(k, p) for p in params] for k in ("__bcmi_radix_sort", "__bcmi_radix_sort_by_key")}
    hparams = ("unroll", "subs", "blocks_per_cu", "bank_groups")
    hkeys = ("__bcmi_radix_hist", "__bcmi_radix_hist_by_key")
    hsaved = {hk: [_tune_get(hk, p) for p in hparams] for hk in hkeys}
    hshape = {(256, 16): (1, 8, 4, 1), (256, 32): (2, 16, 2, 0), (256, 64): (8, 8, 2, 1), (512, 16): (4, 16, 4, 1),
              (512, 24): (8, 16, 2, 0), (512, 32): (4, 16, 2, 1), (1024, 16): (2, 8, 8, 0)}[(threads, ipt)]
    xrun = {(256, 16): 2, (256, 32): 8, (256, 64): 4, (512, 16): 8, (512, 24): 4, (512, 32): 2,
            (1024, 16): 8}[(threads, ipt)]
    if split > 1:
        xrun = 8 // split
    L = ca.lib()
    try:
        for key in saved:
            ca.check(L.bcmi_tune_set(key.encode(), b"threads", threads))
            ca.check(L.bcmi_tune_set(key.encode(), b"ipt", ipt))
            ca.check(L.bcmi_tune_set(key.encode(), b"fit_size", 0))  # the variant at every size
            ca.check(L.bcmi_tune_set(key.encode(), b"xcd_run", xrun))
            ca.check(L.bcmi_tune_set(key.encode(), b"split", split))
        for hk in hkeys:
            for p, val in zip(hparams, hshape):
                ca.check(L.bcmi_tune_set(hk.encode(), p.encode(), val))
        for n in (70001, (1 << 20) + 333, 5000011):
            k = random_keys("uint32", n)
            d = dev(queue, k)
            ca.sort(queue, d, api=ca.RADIX_SORT)
            assert np.array_equal(d.to_host(queue), np.sort(k)), n
            kf = random_keys("float32", n)
            v = np.arange(n, dtype=np.uint32)
            for desc in (False, True):
                dk, dv = dev(queue, kf), dev(queue, v)
                ca.sort(queue, dk, dv, api=ca.SORT_BY_KEY, descending=desc)
                ek, ev = O.radix_order_sort(kf, v, descending=desc)
                assert np.array_equal(bits(dk.to_host(queue)), bits(ek)), (n, desc)
                assert np.array_equal(dv.to_host(queue), ev), (n, desc)
    finally:
        for key, vals in saved.items():
            for p, val in zip(params, vals):
                # an absent entry reads back as the sentinel: restore the built-in default
                ca.check(L.bcmi_tune_set(key.encode(), p.encode(), val if val != 0xFFFFFFFF else defaults[p]))
        for hk in hkeys:
            for p, val, dflt in zip(hparams, hsaved[hk], (4, 16, 2, 0)):
                if val != 0xFFFFFFFF or hk == "__bcmi_radix_hist":
                    ca.check(L.bcmi_tune_set(hk.encode(), p.encode(), val if val != 0xFFFFFFFF else dflt))


@pytest.mark.parametrize("threads,ipt,split,lb,waves,cbg", [
    (256, 64, 2, 2, 0, 0), (512, 32, 2, 2, 0, 0), (256, 64, 2, 2, 4, 0), (512, 24, 2, 2, 8, 0), (512, 32, 2, 2, 7, 0),
    (512, 16, 1, 4, 8, 0), (512, 64, 2, 2, 4, 0), (512, 64, 4, 2, 4, 0), (1024, 32, 2, 2, 0, 0), (1024, 32, 4, 2, 0, 0),
    (512, 64, 2, 2, 4, 1), (256, 64, 2, 2, 4, 1), (512, 32, 2, 2, 0, 1), (512, 32, 2, 2, 7, 1)])
def test_sort_lookback_and_register_bound_variants(queue, threads, ipt, split, lb, waves, cbg):
    """The look-back windows, register bounds and count-word layouts the
    tuner may select: keys and pairs, many tiles (the walk crosses windows),
    XCD-grouped."""
    L = ca.lib()
    params = ("threads", "ipt", "split", "lookback", "waves", "xcd_run", "fit_size", "count_bg")
    keys = ("__bcmi_radix_sort", "__bcmi_radix_sort_by_key")
    saved = {k: [_tune_get(k, p) for p in params] for k in keys}
    defaults = {"threads": 512, "ipt": 32, "split": 1, "lookback": 4, "waves": 0, "xcd_run": 0, "fit_size": 1,
                "count_bg": 0}
    try:
        for key in keys:
            for p, val in zip(params, (threads, ipt, split, lb, waves, 8, 0, cbg)):
                ca.check(L.bcmi_tune_set(key.encode(), p.encode(), val))
        for n in (70001, 5000011):
            k = random_keys("uint32", n)
            d = dev(queue, k)
            ca.sort(queue, d, api=ca.RADIX_SORT)
            assert np.array_equal(d.to_host(queue), np.sort(k)), n
            kf = random_keys("float32", n)
            v = np.arange(n, dtype=np.uint32)
            dk, dv = dev(queue, kf), dev(queue, v)
            ca.sort(queue, dk, dv, api=ca.SORT_BY_KEY)
            ek, ev = O.radix_order_sort(kf, v)
            assert np.array_equal(bits(dk.to_host(queue)), bits(ek)), n
            assert np.array_equal(dv.to_host(queue), ev), n
    finally:
        for key, vals in saved.items():
            for p, val in zip(params, vals):
                ca.check(L.bcmi_tune_set(key.encode(), p.encode(), val if val != 0xFFFFFFFF else defaults[p]))


# ========================================================= segmented sort
def _segments(lens):
    return np.r_[0, np.cumsum(lens)].astype(np.uint64)


@pytest.mark.parametrize("dt", ["uint64", "uint32", "int32", "float32", "float64", "uint8", "int16"])
@pytest.mark.parametrize("vdt", [None, "uint32", "uint64"])
def test_segmented_sort_matches_per_segment_calls(queue, dt, vdt):
    # every size class: empty, singletons, < 32, LDS classes up to 4096, and
    # multi-tile segments for the batched onesweep path
    lens = [0, 1, 2, 5, 31, 32, 33, 200, 256, 257, 700, 1500, 2049, 4096, 4097, 9000, 30000, 0, 3]
    lens += list(RNG.integers(1, 3000, 60))
    off = _segments(lens)
    n = int(off[-1])
    k = random_keys(dt, n)
    v = np.arange(n).astype(vdt) if vdt else None
    for desc in (False, True):
        dk = dev(queue, k)
        dv = dev(queue, v) if vdt else None
        ca.segmented_sort(queue, dk, dv, off, descending=desc)
        exp = O.segmented_sort(k, v, off, descending=desc)
        if vdt:
            ek, ev = exp
            assert np.array_equal(dv.to_host(queue), ev), desc
        else:
            ek = exp
        assert np.array_equal(bits(dk.to_host(queue)), bits(ek)), desc


@pytest.mark.parametrize("waves,q", [(4, 4), (4, 8), (4, 16), (8, 8), (8, 16)])
def test_scan_tunable_shapes(queue, waves, q):
    """Every (waves per workgroup, chunks per lane) shape the tuner may pick
    for the 4-byte scans: integer scans exact (exclusive with an init, in
    place, unaligned), the compensated float scan in the striped layout within
    its bound with both accumulator forms, whole and partial tiles."""
    L = ca.lib()
    keys_t = ("__bcmi_scan", "__bcmi_scan_f32")
    params = ("waves", "q", "acc", "blocked")
    saved = {k: [_tune_get(k, p) for p in params] for k in keys_t}
    try:
        for k in keys_t:
            ca.check(L.bcmi_tune_set(k.encode(), b"waves", waves))
            ca.check(L.bcmi_tune_set(k.encode(), b"q", q))
        ca.check(L.bcmi_tune_set(b"__bcmi_scan_f32", b"blocked", 0))  # the striped layout
        for n in (1, 5000, 65536 * 3 + 7, (1 << 21) + 13):
            x = RNG.integers(-1000, 1000, n).astype(np.int32)
            d = dev(queue, x)
            out = ca.DeviceArray(n, np.int32, queue)
            ca.scan(queue, d, out, True, init=7)
            exp = np.empty(n, np.int64)
            exp[0] = 7
            exp[1:] = 7 + np.cumsum(x[:-1], dtype=np.int64)
            assert np.array_equal(out.to_host(queue), exp.astype(np.int32)), n
            ca.scan(queue, d, d, False)
            assert np.array_equal(d.to_host(queue), np.cumsum(x, dtype=np.int64).astype(np.int32)), n
            if n > 2:
                # input at +1 element, output at +2: neither 16-byte aligned
                d = dev(queue, x)
                m = n - 2
                ca.check(L.bcmi_scan(queue.h, ca.I32, ca.PLUS, d.at(1), out.at(2), m, None, 0, 0))
                assert np.array_equal(out.to_host(queue)[2:], O.scan(x[1:1 + m], False)), n
            xf = O.gen_lognormal_f32(n)
            fo = ca.DeviceArray(n, np.float32, queue)
            for acc in (0, 1):
                ca.check(L.bcmi_tune_set(b"__bcmi_scan_f32", b"acc", acc))
                ca.scan(queue, dev(queue, xf), fo, False)
                dd = O.scan_f32_dd(xf, False)
                assert np.all(np.abs(fo.to_host(queue).astype(np.float64) - dd) <= _scan_err_bound(xf, dd)), (n, acc)
    finally:
        for k, vals_ in saved.items():
            for p, val, d in zip(params, vals_, (4, 16, 1, 0)):
                ca.check(L.bcmi_tune_set(k.encode(), p.encode(), val if val != 0xFFFFFFFF else d))


@pytest.mark.parametrize("waves,q,blocked", [(4, 8, 1), (4, 16, 1), (8, 8, 1),
                                             (4, 8, 2), (4, 16, 2), (8, 8, 2), (8, 16, 2)])
def test_scan_blocked_layout(queue, waves, q, blocked):
    """The compensated float scan with lanes scanning consecutive elements
    (tuning param "blocked": 1 = one transpose round, 2 = two rounds of half
    the chunks), both accumulator forms: inclusive, exclusive with an init,
    in place and unaligned, whole and partial tiles, within the bound of the
    double-double oracle."""
    L = ca.lib()
    key = "__bcmi_scan_f32"
    params = ("waves", "q", "acc", "blocked")
    saved = [_tune_get(key, p) for p in params]
    try:
        ca.check(L.bcmi_tune_set(key.encode(), b"waves", waves))
        ca.check(L.bcmi_tune_set(key.encode(), b"q", q))
        ca.check(L.bcmi_tune_set(key.encode(), b"blocked", blocked))
        for n in (1, 77, 5000, 65536 * 3 + 7, (1 << 21) + 13):
            x = O.gen_lognormal_f32(n)
            x[::3] *= -1.0  # mixed signs: cancellation inside and across tiles
            fo = ca.DeviceArray(n, np.float32, queue)
            for acc in (0, 1):
                ca.check(L.bcmi_tune_set(key.encode(), b"acc", acc))
                ca.scan(queue, dev(queue, x), fo, False)
                dd = O.scan_f32_dd(x, False)
                got = fo.to_host(queue).astype(np.float64)
                assert np.all(np.abs(got - dd) <= _scan_err_bound(x, dd)), (n, acc, "inclusive")
                ca.scan(queue, dev(queue, x), fo, True, init=np.float32(2.5))
                dd = O.scan_f32_dd(x, True, 2.5)
                got = fo.to_host(queue).astype(np.float64)
                assert np.all(np.abs(got - dd) <= _scan_err_bound(x, dd) + np.spacing(np.float32(2.5))), \
                    (n, acc, "exclusive")
                d = dev(queue, x)
                ca.scan(queue, d, d, False)
                dd = O.scan_f32_dd(x, False)
                assert np.all(np.abs(d.to_host(queue).astype(np.float64) - dd) <= _scan_err_bound(x, dd)), \
                    (n, acc, "in place")
                if n > 2:
                    d = dev(queue, x)
                    m = n - 2
                    ca.check(L.bcmi_scan(queue.h, ca.F32, ca.PLUS, d.at(1), fo.at(2), m, None, 0, ca.COMPENSATED))
                    dd = O.scan_f32_dd(x[1:1 + m], False)
                    got = fo.to_host(queue)[2:].astype(np.float64)
                    assert np.all(np.abs(got - dd) <= _scan_err_bound(x[1:1 + m], dd)), (n, acc, "unaligned")
    finally:
        for p, val, d in zip(params, saved, (4, 16, 1, 0)):
            ca.check(L.bcmi_tune_set(key.encode(), p.encode(), val if val != 0xFFFFFFFF else d))


@pytest.mark.parametrize("waves,q", [(4, 2), (4, 4), (4, 8), (8, 8)])
def test_segmented_scan_tunable_shapes(queue, waves, q):
    """Every (waves per workgroup, chunks per lane) shape the tuner may pick
    for reduce_by_key and the segmented scans: integer results exact, float
    sums within the compensated bound, whole and partial tiles."""
    L = ca.lib()
    keys_t = ("__bcmi_reduce_by_key", "__bcmi_segscan")
    params = ("waves", "q")
    saved = {k: [_tune_get(k, p) for p in params] for k in keys_t}
    try:
        for k in keys_t:
            ca.check(L.bcmi_tune_set(k.encode(), b"waves", waves))
            ca.check(L.bcmi_tune_set(k.encode(), b"q", q))
        for n in (1, 4097, 100003, (1 << 20) + 17):
            lens = RNG.integers(1, 70, n)
            keys = np.repeat(np.arange(lens.size, dtype=np.int32), lens)[:n]
            vals = RNG.integers(-100, 100, n).astype(np.int32)
            ko = ca.DeviceArray(n, np.int32, queue)
            vo = ca.DeviceArray(n, np.int32, queue)
            r = ca.reduce_by_key(queue, dev(queue, keys), dev(queue, vals), ko, vo)
            ek, ev = O.reduce_by_key(keys, vals)
            assert r == ek.size and np.array_equal(ko.to_host(queue)[:r], ek) and \
                np.array_equal(vo.to_host(queue)[:r], ev), n
            for excl in (False, True):
                ca.segmented_scan(queue, dev(queue, keys), dev(queue, vals), vo, excl, init=4 if excl else None)
                assert np.array_equal(vo.to_host(queue), O.segmented_scan(keys, vals, excl, 4 if excl else 0)), n
            fk, fv = O.gen_runs_i32_f32(n)
            fo = ca.DeviceArray(n, np.float32, queue)
            r = ca.reduce_by_key(queue, dev(queue, fk), dev(queue, fv), ko, fo)
            ek, es = O.reduce_by_key_f32_dd(fk, fv)
            assert r == ek.size and np.array_equal(ko.to_host(queue)[:r], ek), n
            absum = np.add.reduceat(np.abs(fv.astype(np.float64)), np.r_[0, np.flatnonzero(np.diff(fk)) + 1])
            bound = np.spacing(np.abs(es).astype(np.float32)).astype(np.float64) + 2.0 ** -50 * absum
            assert np.all(np.abs(fo.to_host(queue)[:r].astype(np.float64) - es) <= bound), n
            ca.segmented_scan(queue, dev(queue, fk), dev(queue, fv), fo, True, init=0.0)
            dd = O.segmented_scan_f32_dd(fk, fv, True)
            bound = np.spacing(np.abs(dd).astype(np.float32)).astype(np.float64) + \
                2.0 ** -50 * np.cumsum(np.abs(fv.astype(np.float64)))
            assert np.all(np.abs(fo.to_host(queue).astype(np.float64) - dd) <= bound), n
    finally:
        for k, vals_ in saved.items():
            for p, val, d in zip(params, vals_, (4, 8)):
                ca.check(L.bcmi_tune_set(k.encode(), p.encode(), val if val != 0xFFFFFFFF else d))


@pytest.mark.parametrize("acc", [0, 1])
def test_float_accumulator_policies(queue, acc):
    """Both float accumulator policies the tuner may pick (1: FP64
    accumulators, 0: TwoSum FP32 pairs) meet the compensated bound for the
    scan, reduce_by_key and the segmented scan, on whole and partial tiles."""
    L = ca.lib()
    keys_t = ("__bcmi_scan_f32", "__bcmi_reduce_by_key", "__bcmi_segscan")
    saved = [_tune_get(k, "acc") for k in keys_t]
    try:
        for k in keys_t:
            ca.check(L.bcmi_tune_set(k.encode(), b"acc", acc))
        for n in (4096 * 3 + 5, (1 << 20) + 17):
            x = O.gen_lognormal_f32(n)
            out = ca.DeviceArray(n, np.float32, queue)
            ca.scan(queue, dev(queue, x), out, False)
            dd = O.scan_f32_dd(x, False)
            assert np.all(np.abs(out.to_host(queue).astype(np.float64) - dd) <= _scan_err_bound(x, dd)), n
            keys, vals = O.gen_runs_i32_f32(n)
            ko = ca.DeviceArray(n, np.int32, queue)
            r = ca.reduce_by_key(queue, dev(queue, keys), dev(queue, vals), ko, out)
            ek, es = O.reduce_by_key_f32_dd(keys, vals)
            assert r == ek.size and np.array_equal(ko.to_host(queue)[:r], ek), n
            absum = np.add.reduceat(np.abs(vals.astype(np.float64)), np.r_[0, np.flatnonzero(np.diff(keys)) + 1])
            bound = np.spacing(np.abs(es).astype(np.float32)).astype(np.float64) + 2.0 ** -50 * absum
            assert np.all(np.abs(out.to_host(queue)[:r].astype(np.float64) - es) <= bound), n
            for excl in (False, True):
                ca.segmented_scan(queue, dev(queue, keys), dev(queue, vals), out, excl, init=0.0 if excl else None)
                dd = O.segmented_scan_f32_dd(keys, vals, excl)
                bound = np.spacing(np.abs(dd).astype(np.float32)).astype(np.float64) + \
                    2.0 ** -50 * np.cumsum(np.abs(vals.astype(np.float64)))
                assert np.all(np.abs(out.to_host(queue).astype(np.float64) - dd) <= bound), (n, excl)
    finally:
        for k, val in zip(keys_t, saved):
            ca.check(L.bcmi_tune_set(k.encode(), b"acc", val if val != 0xFFFFFFFF else 1))


@pytest.mark.parametrize("small_max,msd,ipt,qrank", [(2, 1, 12, 0), (3, 0, 12, 0), (0, 1, 12, 0), (3, 1, 24, 0),
                                                      (2, 0, 24, 0), (3, 1, 24, 1), (3, 0, 24, 1), (2, 1, 12, 1)])
def test_segmented_sort_routing_thresholds(queue, small_max, msd, ipt, qrank):
    """Every routing threshold, large-tile size and large-pass ranking the
    tuner may pick (small classes capped, MSD route off, 12 or 24 keys per
    thread, whole-wave or quarter-wave ranking) gives the per-segment result: segments between the cap and 4096 go to the large
    path and, with the MSD route, to its groups."""
    L = ca.lib()
    saved = [_tune_get("__bcmi_segsort", p) for p in ("small_max", "msd", "ipt", "qrank")]
    try:
        ca.check(L.bcmi_tune_set(b"__bcmi_segsort", b"small_max", small_max))
        ca.check(L.bcmi_tune_set(b"__bcmi_segsort", b"msd", msd))
        ca.check(L.bcmi_tune_set(b"__bcmi_segsort", b"ipt", ipt))
        ca.check(L.bcmi_tune_set(b"__bcmi_segsort", b"qrank", qrank))
        lens = [0, 1, 2, 31, 32, 200, 256, 257, 700, 1025, 2049, 4096, 4097, 9000, 30000]
        lens += list(RNG.integers(1, 5000, 40))
        off = _segments(lens)
        n = int(off[-1])
        for dt in ("uint64", "float32"):
            k = random_keys(dt, n)
            v = np.arange(n, dtype=np.uint32)
            dk, dv = dev(queue, k), dev(queue, v)
            plan = ca.SegmentedPlan(queue, np.dtype(dt), np.uint32, off)
            plan.run(queue, dk, dv)
            ek, ev = O.segmented_sort(k, v, off)
            assert np.array_equal(bits(dk.to_host(queue)), bits(ek)), dt
            assert np.array_equal(dv.to_host(queue), ev), dt
    finally:
        for p, val, d in zip(("small_max", "msd", "ipt", "qrank"), saved, (4, 1, 12, 0)):
            ca.check(L.bcmi_tune_set(b"__bcmi_segsort", p.encode(), val if val != 0xFFFFFFFF else d))


def test_segmented_sort_float_small_segments_keep_signed_zero_order(queue):
    # below 32 elements stable_sort_by_key is the insertion sort: +0.0 and
    # -0.0 compare equal and keep input order; from 32 on, radix order puts
    # -0.0 first
    k = np.array([0.0, -0.0, 1.0] * 11, np.float32)  # 33 keys
    off = np.array([0, 3, 33], np.uint64)
    v = np.arange(33, dtype=np.uint32)
    dk, dv = dev(queue, k), dev(queue, v)
    ca.segmented_sort(queue, dk, dv, off)
    ek, ev = O.segmented_sort(k, v, off)
    assert np.array_equal(dv.to_host(queue), ev)
    assert list(ev[:3]) == [0, 1, 2]


@pytest.mark.parametrize("dt", ["uint32", "uint64", "float32", "int16"])
def test_segmented_sort_routes_follow_the_data(queue, dt):
    """Large segments whose top digit splits them into buckets that fit a
    group take the top-digit route; a segment whose keys share the top digit
    (one bucket larger than a group) falls back to all digit passes.  The
    route is chosen on the device per run, so one plan run on two inputs
    with the roles swapped must match the per-segment sorts both times."""
    lens = [20000, 20000, 50, 9000, 0, 4097]
    off = _segments(lens)
    n = int(off[-1])
    uni = random_keys(dt, n)
    ubits = np.dtype("u%d" % uni.itemsize)
    low = np.iinfo(ubits).max >> 8

    def one_bucket(sl):  # clear the top byte of the keys in sl
        k = uni.copy()
        k.view(ubits)[sl] &= low
        return k

    v = np.arange(n, dtype=np.uint32)
    plan = ca.SegmentedPlan(queue, np.dtype(dt), np.uint32, off)
    for keys in (one_bucket(slice(0, 20000)), one_bucket(slice(20000, 40000))):
        dk, dv = dev(queue, keys), dev(queue, v)
        plan.run(queue, dk, dv)
        ek, ev = O.segmented_sort(keys, v, off)
        assert np.array_equal(bits(dk.to_host(queue)), bits(ek))
        assert np.array_equal(dv.to_host(queue), ev)


@pytest.mark.parametrize("dt", ["uint64", "int32", "float64", "uint16"])
def test_segmented_sort_in_lds_paths(queue, dt):
    """Spans sorted in LDS take one range-bucketing pass plus per-bucket
    insertion sorts when the keys spread over their range, all LSD passes
    when they cluster (a bucket over 64 keys), and a plain copy when every
    key is equal; each must match the per-segment sort."""
    rng = np.random.default_rng(7)
    spread = random_keys(dt, 3000, rng)
    clustered = np.repeat(random_keys(dt, 4, rng), 750)
    clustered[rng.integers(0, 3000, 40)] = random_keys(dt, 40, rng)
    equal = np.full(2500, spread[0])
    keys = np.concatenate([spread, clustered, equal, spread[:700], clustered[:100]]).astype(dt)
    off = _segments([3000, 3000, 2500, 700, 100])
    v = np.arange(keys.size, dtype=np.uint32)
    for desc in (False, True):
        dk, dv = dev(queue, keys), dev(queue, v)
        ca.segmented_sort(queue, dk, dv, off, descending=desc)
        ek, ev = O.segmented_sort(keys, v, off, descending=desc)
        assert np.array_equal(bits(dk.to_host(queue)), bits(ek)), desc
        assert np.array_equal(dv.to_host(queue), ev), desc


def test_segmented_sort_workload_sample(queue):
    """The segmented workload's generator at 2048 segments (lognormal lengths,
    median 256, one 2^20 segment): bit-identical to per-segment sorts."""
    off = O.gen_segments(2048)
    n = int(off[-1])
    k, v = O.gen_segment_data(n)
    dk, dv = dev(queue, k), dev(queue, v)
    ca.segmented_sort(queue, dk, dv, off)
    ek, ev = O.segmented_sort(k, v, off)
    assert np.array_equal(dk.to_host(queue), ek)
    assert np.array_equal(dv.to_host(queue), ev)


def test_segmented_plan_graph_replay(queue):
    lens = list(RNG.integers(1, 6000, 300)) + [50000]
    off = _segments(lens)
    n = int(off[-1])
    k = random_keys("uint64", n)
    v = np.arange(n, dtype=np.uint32)
    dk, dv = dev(queue, k), dev(queue, v)
    plan = ca.SegmentedPlan(queue, np.uint64, np.uint32, off)
    plan.run(queue, dk, dv)  # eager
    ek, ev = O.segmented_sort(k, v, off)
    assert np.array_equal(dk.to_host(queue), ek) and np.array_equal(dv.to_host(queue), ev)
    g = ca.Graph(queue)
    with g:
        plan.run(queue, dk, dv)
    for rep in range(3):
        ca.check(ca.lib().bcmi_memcpy(queue.h, dk.ptr, k.ctypes.data_as(ca.C.c_void_p), k.nbytes, ca.H2D, 1))
        ca.check(ca.lib().bcmi_memcpy(queue.h, dv.ptr, v.ctypes.data_as(ca.C.c_void_p), v.nbytes, ca.H2D, 1))
        g.launch()
        queue.finish()
        assert np.array_equal(dk.to_host(queue), ek), rep
        assert np.array_equal(dv.to_host(queue), ev), rep
        x = dev(queue, np.arange(1000, dtype=np.int32))  # eager work between replays
        ca.scan(queue, x, x, False)
        queue.finish()
