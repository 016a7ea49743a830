"""Parity of transform_reduce / inner_product (bcmi_transform_reduce) with the
oracle and the reference's golden vectors.  Bar: bit-exact for integers and
for float min/max (order-free); float sums within 2 ulp of the exact sum of
exact products plus 2^-30 of the sum of |terms| (the cancellation term of a
compensated dot product)."""
import math

import numpy as np
import pytest

import compute_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(777)
INTS = ["int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64"]
ROPS = [(ca.PLUS, O.PLUS), (ca.MULTIPLIES, O.MULTIPLIES), (ca.MIN, O.MIN), (ca.MAX, O.MAX)]
MAPS = [("multiplies", "id", 0), ("multiplies", "abs", 0), ("multiplies", "square", 0), ("minus", "abs", 0),
        ("minus", "square", 0), ("plus", "id", 0), (None, "abs", 0), (None, "square", 3), (None, "id", 5)]


def dev(q, a):
    return ca.DeviceArray.from_host(np.ascontiguousarray(a), q)


def run(q, dt, a, b, n, op, fn, c, rop, init=None, precision=ca.COMPENSATED):
    out = dev(q, np.zeros(1, dt))
    f64 = ca.transform_reduce(q, dt, a, b, n, ca.mapping(dt, fn, op or "plus", c), out, rop, init, precision)
    return out.to_host(q)[0], float(f64.to_host(q)[0])


def test_golden_transform_reduce(queue, golden):
    seen = set()
    for c in golden:
        if c["algo"] not in ("transform_reduce", "inner_product"):
            continue
        seen.add(c["algo"])
        dt = np.dtype(c["dtype"])

        def seq(spec):
            return ca.counting(spec["count"]) if isinstance(spec, dict) else dev(queue, np.array(spec, dt))
        n = c["a"]["n"] if isinstance(c["a"], dict) else len(c["a"])
        if c["algo"] == "inner_product":
            got, _ = run(queue, dt, seq(c["a"]), seq(c["b"]), n, "multiplies", "id", 0, ca.PLUS, c["init"])
            assert got == c["expected"], c["id"]
        else:
            got, _ = run(queue, dt, seq(c["a"]), None, n, None, c["map"]["fn"], c["map"]["c"], ca.PLUS)
            if dt.kind == "f":
                assert abs(got - c["expected"]) <= c["rtol"] * abs(c["expected"]), c["id"]
            else:
                assert got == c["expected"], c["id"]
    assert seen == {"transform_reduce", "inner_product"}


@pytest.mark.parametrize("dt", INTS)
def test_integer_maps_exact(queue, dt):
    info = np.iinfo(dt)
    for n in (1, 7, 64, 1000, 4099, (1 << 18) + 5):
        a = RNG.integers(info.min, info.max, n, endpoint=True, dtype=dt)
        b = RNG.integers(info.min, info.max, n, endpoint=True, dtype=dt)
        da, db = dev(queue, a), dev(queue, b)
        for op, fn, c in MAPS:
            m = O.mapping(dt, fn, op or "plus", c)
            for (crop, orop) in ROPS:
                got, _ = run(queue, dt, da, db if op else None, n, op, fn, c, crop)
                want = O.map_reduce(dt, a, b if op else None, n, m, orop)
                assert got == want, (n, op, fn, c, crop)
        # init folded first (accumulate / inner_product)
        got, _ = run(queue, dt, da, db, n, "multiplies", "id", 0, ca.PLUS, init=9)
        assert got == O.map_reduce(dt, a, b, n, O.mapping(dt), O.PLUS, 9)


@pytest.mark.parametrize("dt", ["int32", "uint64", "float32", "float64"])
def test_subranges_and_counting_sequences(queue, dt):
    n = 50001
    a = (RNG.integers(-1000, 1000, n + 8)).astype(dt)
    b = (RNG.integers(-1000, 1000, n + 8)).astype(dt)
    da, db = dev(queue, a), dev(queue, b)
    m = O.mapping(dt)
    for oa, ob in ((1, 1), (1, 2), (3, 0), (0, 5)):  # same and different offsets inside a 16-byte chunk
        sa, sb = ca._as_seq(da, oa), ca._as_seq(db, ob)
        got, _ = run(queue, dt, sa, sb, n, "multiplies", "id", 0, ca.PLUS)
        xa, xb = a[oa:oa + n], b[ob:ob + n]
        if np.dtype(dt).kind == "f":
            hi, lo = O.map_sum_dd(dt, xa, xb, n, m)
            # integer-valued terms: float64 sums them exactly, float32 rounds once
            assert abs(float(got) - (hi + lo)) <= float(np.spacing(np.array(abs(hi + lo), dt))), (oa, ob)
        else:
            assert got == O.map_reduce(dt, xa, xb, n, m, O.PLUS), (oa, ob)
    # counting sequences, alone and mixed with an array
    got, _ = run(queue, dt, ca.counting(-7), ca.counting(3), 20000, "multiplies", "id", 0, ca.PLUS)
    if np.dtype(dt).kind == "f":
        hi, lo = O.map_sum_dd(dt, ("count", -7), ("count", 3), 20000, m)
        assert abs(float(got) - (hi + lo)) <= float(np.spacing(np.array(abs(hi + lo), dt)))
    else:
        assert got == O.map_reduce(dt, ("count", -7), ("count", 3), 20000, m, O.PLUS)
    got, _ = run(queue, dt, da, ca.counting(0), n, "minus", "abs", 0, ca.MAX)
    assert got == O.map_reduce(dt, a, ("count", 0), n, O.mapping(dt, "abs", "minus"), O.MAX)


def _bound(dt, exact, terms_abs):
    ulp = np.spacing(np.array(abs(exact), dt)).astype(np.float64)
    return 2 * float(ulp) + 2.0 ** -30 * terms_abs


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("n", [1, 33, 4096, 100003, 1 << 20])
def test_float_sums_compensated(queue, dt, n):
    a = (RNG.normal(size=n) * 10.0 ** RNG.integers(-3, 4, n)).astype(dt)
    b = (RNG.normal(size=n) * 10.0 ** RNG.integers(-3, 4, n)).astype(dt)
    da, db = dev(queue, a), dev(queue, b)
    for op, fn, c in [("multiplies", "id", 0), ("minus", "square", 0), (None, "square", 0.25), (None, "abs", 0),
                      ("plus", "id", 0)]:
        m = O.mapping(dt, fn, op or "plus", c)
        hi, lo = O.map_sum_dd(dt, a, b if op else None, n, m)
        exact = hi + lo
        got, got64 = run(queue, dt, da, db if op else None, n, op, fn, c, ca.PLUS)
        # the size of the terms, for the cancellation part of the bound
        if op == "multiplies":
            terms = O.map_sum_dd(dt, np.abs(a), np.abs(b), n, O.mapping(dt))[0]
        elif op == "plus":
            terms = float(np.sum(np.abs(a), dtype=np.float64) + np.sum(np.abs(b), dtype=np.float64))
        else:
            terms = abs(exact)  # squares and absolute values: no cancellation
        assert abs(float(got) - exact) <= _bound(dt, exact, terms), (op, fn, c, float(got), exact)
        assert abs(got64 - exact) <= 2.0 ** -40 * max(terms, abs(exact)) + 1e-300, (op, fn, c)


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_float_min_max_exact(queue, dt):
    n = 100003
    a = RNG.normal(size=n).astype(dt)
    b = RNG.normal(size=n).astype(dt)
    da, db = dev(queue, a), dev(queue, b)
    for op, fn, c in [("multiplies", "id", 0), ("minus", "abs", 0), (None, "square", 0.5)]:
        m = O.mapping(dt, fn, op or "plus", c)
        for crop, orop in ((ca.MIN, O.MIN), (ca.MAX, O.MAX)):
            got, _ = run(queue, dt, da, db if op else None, n, op, fn, c, crop)
            assert got == O.map_reduce(dt, a, b if op else None, n, m, orop), (op, fn, crop)


def test_naive_float_sum_is_the_plain_fp32_fold(queue):
    n = 1 << 16
    a = RNG.normal(size=n).astype(np.float32)
    b = RNG.normal(size=n).astype(np.float32)
    got, _ = run(queue, np.float32, dev(queue, a), dev(queue, b), n, "multiplies", "id", 0, ca.PLUS,
                 precision=ca.NAIVE)
    exact = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    assert abs(float(got) - exact) <= 1e-4 * float(np.dot(np.abs(a), np.abs(b)))


def test_empty_range(queue):
    a = dev(queue, np.arange(4, dtype=np.int32))
    out = dev(queue, np.array([42], np.int32))
    ca.transform_reduce(queue, np.int32, a, a, 0, ca.mapping(np.int32), out)
    assert out.to_host(queue)[0] == 42  # reduce of an empty range leaves the result
    ca.transform_reduce(queue, np.int32, a, a, 0, ca.mapping(np.int32), out, init=5)
    assert out.to_host(queue)[0] == 5   # accumulate / inner_product of an empty range is init


def test_inner_product_full_size(queue):
    # 2^28 float32 pairs: the compensated dot product against float64 numpy
    n = 1 << 28
    x, y = O.gen_saxpy_f32(n)
    got, got64 = run(queue, np.float32, dev(queue, x), dev(queue, y), n, "multiplies", "id", 0, ca.PLUS)
    ref = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
    scale = float(np.dot(np.abs(x).astype(np.float64), np.abs(y).astype(np.float64)))
    assert abs(got64 - ref) <= 1e-9 * scale
    assert abs(float(got) - ref) <= 2 * float(np.spacing(np.float32(abs(ref)))) + 1e-9 * scale
