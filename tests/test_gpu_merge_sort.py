"""Parity of the comparator merge sort (bcmi_merge_sort: sort / stable_sort /
sort_by_key / stable_sort_by_key with a user comparator) with the oracle and
the reference's golden vectors.  Bar: bit-exact records and values (the
stable order) at every size, dtype, key form and direction."""
import numpy as np
import pytest

import compute_amd as ca
import oracle as O
from test_oracle import CMP_DTYPES, _cmp_records, check_compare_case, compare_case_arrays, compare_case_cmp

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=[1, 0], ids=["key_route", "merge_kernels"])
def route(request):
    """Every test runs on both routes of a comparator that reads the same field
    of both operands: the radix sort of its key codes (the default) and the
    merge kernels.  Asymmetric comparators always take the merge kernels."""
    ca.check(ca.lib().bcmi_tune_set(b"__bcmi_merge_sort", b"route", request.param))
    yield request.param
    ca.check(ca.lib().bcmi_tune_set(b"__bcmi_merge_sort", b"route", 1))

# tile sizes of the kernels: 4096 items for 8-byte items (fields of <= 4
# bytes), 2048 for 16- and 24-byte items
EDGE_SIZES = [2, 3, 31, 32, 33, 2047, 2048, 2049, 4095, 4096, 4097, 3 * 4096 + 5, 50001]


def dev(q, a):
    return ca.DeviceArray.from_host(np.ascontiguousarray(a), q)


def gpu_sort(q, rec, rb, dt, values=None, **kw):
    n = rec.nbytes // rb
    d = dev(q, rec)
    dv = dev(q, values) if values is not None else None
    cmp = ca.comparator(dt, kw.get("form", "field"), kw.get("lhs", 0), kw.get("rhs"), kw.get("second", 0),
                        kw.get("count", 0), kw.get("greater", False))
    vb = values.nbytes // n if values is not None and n else None
    ca.merge_sort(q, cmp, d, rb, n, dv, vb)
    out = d.to_host(q).view(rec.dtype).reshape(rec.shape)
    return out, (dv.to_host(q).view(values.dtype).reshape(values.shape) if values is not None else None)


def assert_same(q, rec, rb, dt, values=None, **kw):
    got, gv = gpu_sort(q, rec, rb, dt, values, **kw)
    want, wv = O.comparator_sort(rec, rb, dt, values=values, **kw)
    assert got.tobytes() == want.tobytes(), kw
    if values is not None:
        assert gv.tobytes() == wv.tobytes(), kw


def test_golden_compare_sort(queue, golden):
    seen = 0
    for c in golden:
        if c["algo"] != "compare_sort":
            continue
        seen += 1
        rec, rb, vals = compare_case_arrays(c)
        kw = compare_case_cmp(c)
        got, gv = gpu_sort(queue, rec, rb, c["dtype"], vals, **kw)
        check_compare_case(c, got.reshape(-1), gv)
        want, wv = O.comparator_sort(rec, rb, c["dtype"], values=vals, **kw)
        assert got.tobytes() == want.tobytes(), c["id"]
    assert seen == 13


@pytest.mark.parametrize("dt", CMP_DTYPES)
def test_all_forms_and_directions(queue, dt):
    rng = np.random.default_rng(2024)
    sz = np.dtype(dt).itemsize
    forms = ["field", "abs", "abs_sum2"] + (["length"] if np.dtype(dt).kind == "f" else [])
    for comps in (1, 2, 4):
        rb = comps * sz
        for form in forms:
            if form == "abs_sum2" and comps == 1:
                continue
            for greater in (False, True):
                for n in (33, 4097, 50001):
                    rec = _cmp_records(dt, n, comps, rng)
                    vals = np.arange(n, dtype=np.uint32)
                    kw = dict(form=form, lhs=sz * (comps - 1 if form == "field" else 0), second=sz * (comps - 1),
                              count=comps, greater=greater)
                    assert_same(queue, rec, rb, dt, vals, **kw)


@pytest.mark.parametrize("dt", ["int8", "int32", "float32", "uint64", "float64"])
def test_tile_boundaries(queue, dt):
    rng = np.random.default_rng(7)
    for n in EDGE_SIZES:
        rec = _cmp_records(dt, n, 2, rng)
        vals = np.arange(n, dtype=np.uint32)
        sz = np.dtype(dt).itemsize
        assert_same(queue, rec, 2 * sz, dt, vals, form="field", lhs=sz)
        assert_same(queue, rec, 2 * sz, dt, None, form="abs", greater=True)


def test_record_and_value_sizes(queue):
    rng = np.random.default_rng(11)
    n = 9001
    for dt, comps in (("uint8", 1), ("int16", 1), ("int32", 1), ("int32", 2), ("float32", 4), ("int64", 4),
                      ("int64", 8), ("float64", 8), ("int8", 16)):
        rec = _cmp_records(dt, n, comps, rng)
        sz = np.dtype(dt).itemsize
        for vb in (1, 2, 4, 8, 12, 16, 24):
            vals = rng.integers(0, 256, (n, vb), dtype=np.uint8)
            assert_same(queue, rec, comps * sz, dt, vals, form="field", lhs=(comps - 1) * sz)


def test_asymmetric_comparator(queue):
    # a.s0 < b.s3 (test/test_merge_sort_gpu.cpp:311-314) on long8_ records
    # with equal components: the order of the components, stable
    rng = np.random.default_rng(5)
    for n in (256, 2049, 100003):
        base = rng.integers(-1000, 1000, n).astype(np.int64)
        rec = np.repeat(base[:, None], 8, axis=1)
        vals = np.arange(n, dtype=np.uint32)
        got, gv = gpu_sort(queue, rec, 64, np.int64, vals, form="field", lhs=0, rhs=24)
        order = np.argsort(base, kind="stable")
        assert np.array_equal(got[:, 0], base[order]) and np.array_equal(gv, vals[order])
    # int32 fields read from different components of int4 records that are
    # equal by construction, both directions
    for greater in (False, True):
        b = rng.integers(-50, 50, 5000).astype(np.int32)
        rec = np.stack([b, b * 0 + 3, b, b * 0 - 7], axis=1)
        assert_same(queue, rec, 16, np.int32, np.arange(5000, dtype=np.uint32), form="abs", lhs=0, rhs=8,
                     greater=greater)


def test_large_stable_property(queue):
    # 2^22 int2_ records by x: sorted, stable (values = input positions
    # increase inside every run of equal x), and bit-exact to the oracle
    n = 1 << 22
    rng = np.random.default_rng(3)
    rec = np.stack([rng.integers(-5000, 5000, n), rng.integers(-9, 9, n)], axis=1).astype(np.int32)
    vals = np.arange(n, dtype=np.uint32)
    got, gv = gpu_sort(queue, rec, 8, np.int32, vals, form="field", lhs=0)
    x = got[:, 0]
    assert np.all(x[1:] >= x[:-1])
    same = x[1:] == x[:-1]
    assert np.all(gv[1:][same] > gv[:-1][same])
    want, wv = O.comparator_sort(rec, 8, np.int32, values=vals, form="field", lhs=0)
    assert got.tobytes() == want.tobytes() and gv.tobytes() == wv.tobytes()


def test_trivial_and_invalid(queue):
    a = dev(queue, np.array([5, 3], np.int32))
    ca.merge_sort(queue, ca.comparator(np.int32), a, 4, 0)
    ca.merge_sort(queue, ca.comparator(np.int32), a, 4, 1)
    assert a.to_host(queue).tolist() == [5, 3]
    ca.merge_sort(queue, ca.comparator(np.int32), a, 4, 2)
    assert a.to_host(queue).tolist() == [3, 5]
    r12 = dev(queue, np.zeros(12, np.int32))
    with pytest.raises(ca.BcmiError):
        ca.merge_sort(queue, ca.comparator(np.int32), r12, 12, 4)  # 12-byte records
    with pytest.raises(ca.BcmiError):
        ca.merge_sort(queue, ca.comparator(np.int32, lhs=8), a, 8, 1)  # field past the record
    with pytest.raises(ca.BcmiError):
        ca.merge_sort(queue, ca.comparator(np.int32, "length", count=2), a, 8, 1)  # length of ints
    with pytest.raises(ca.BcmiError):
        ca.merge_sort(queue, ca.comparator(np.int32, "abs_sum2", lhs=0, rhs=4, second=4), a, 8, 1)


def test_graph_replay(queue):
    # captured once (after an eager call sizes the scratch arena), replayed on
    # new contents with eager sorts between replays; every replay matches
    rng = np.random.default_rng(21)
    n = 5 * 4096 + 123
    d = dev(queue, np.zeros((n, 2), np.int32))
    dv = dev(queue, np.zeros(n, np.uint32))
    other = dev(queue, rng.integers(-9, 9, (777, 2)).astype(np.int32))
    cmp = ca.comparator(np.int32, "abs", 4, greater=True)
    ca.merge_sort(queue, cmp, d, 8, n, dv, 4)
    queue.finish()
    g = ca.Graph(queue)
    with g:
        ca.merge_sort(queue, cmp, d, 8, n, dv, 4)
    for rep in range(3):
        rec = rng.integers(-1000 * (rep + 1), 1000, (n, 2)).astype(np.int32)
        vals = np.arange(n, dtype=np.uint32)[::-1].copy()
        ca.check(ca.lib().bcmi_memcpy(queue.h, d.ptr, rec.ctypes.data_as(ca.C.c_void_p), rec.nbytes, ca.H2D, 1))
        ca.check(ca.lib().bcmi_memcpy(queue.h, dv.ptr, vals.ctypes.data_as(ca.C.c_void_p), vals.nbytes, ca.H2D, 1))
        ca.merge_sort(queue, ca.comparator(np.int32), other, 8, 777)  # eager work between replays
        g.launch()
        queue.finish()
        want, wv = O.comparator_sort(rec, 8, np.int32, values=vals, form="abs", lhs=4, greater=True)
        got = d.to_host(queue).view(np.int32).reshape(n, 2)
        assert got.tobytes() == want.tobytes() and dv.to_host(queue).tobytes() == wv.tobytes(), rep


def test_sub_range_at_unaligned_offsets(queue):
    # sort [3, n - 2) of int2_ records (24-byte offset: 8-byte copy words) with
    # values at a 1-element offset (4-byte words); elements outside stay put
    rng = np.random.default_rng(31)
    n = 3 * 4096 + 17
    rec = rng.integers(-300, 300, (n, 2)).astype(np.int32)
    vals = rng.integers(0, 1 << 30, n + 1).astype(np.uint32)
    d, dv = dev(queue, rec), dev(queue, vals)
    m = n - 5
    cmp = ca.comparator(np.int32, "field", 4)
    ca.check(ca.lib().bcmi_merge_sort(queue.h, ca.C.byref(cmp), d.at(6), 8, dv.at(1), 4, m))  # record 3 = int32 6
    got = d.to_host(queue).view(np.int32).reshape(n, 2)
    gv = dv.to_host(queue)
    want, wv = O.comparator_sort(rec[3:n - 2].copy(), 8, np.int32, values=vals[1:m + 1].copy(), form="field", lhs=4)
    assert got[3:n - 2].tobytes() == want.tobytes() and gv[1:m + 1].tobytes() == wv.tobytes()
    assert np.array_equal(got[:3], rec[:3]) and np.array_equal(got[n - 2:], rec[n - 2:])
    assert gv[0] == vals[0] and np.array_equal(gv[m + 1:], vals[m + 1:])


@pytest.mark.parametrize("cols", [2, 4])
def test_records_at_misaligned_offset_no_values(queue, cols):
    # int2 / int4 records starting 4 bytes into the buffer (aligned to the
    # int32 field, not to the 8- or 16-byte record) with no values: the key
    # route must not move them as 8- or 16-byte words
    rng = np.random.default_rng(40 + cols)
    n = 2 * 4096 + 77
    flat = rng.integers(-500, 500, n * cols + 1).astype(np.int32)
    d = dev(queue, flat)
    rb = 4 * cols
    cmp = ca.comparator(np.int32, "field", 4)
    ca.check(ca.lib().bcmi_merge_sort(queue.h, ca.C.byref(cmp), d.at(1), rb, None, 0, n))
    got = d.to_host(queue)
    rec = flat[1:].reshape(n, cols).copy()
    want, _ = O.comparator_sort(rec, rb, np.int32, form="field", lhs=4)
    assert got[1:].tobytes() == want.tobytes()
    assert got[0] == flat[0]
