"""Parity of the selection kernels (copy_if / copy_index_if / remove /
count_if / unique / partition_copy / stable_partition) through the C-ABI
with the oracle and the reference's golden vectors.  Bar: bit-exact (these
paths move elements; nothing is rounded)."""
import numpy as np
import pytest

import compute_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(4242)
ALL = ["int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64", "float32", "float64"]
# sizes around the tile boundaries (16384 elements per tile, 8192 for 8-byte
# types) and the staging rounds inside a tile (every 8192 positions)
SIZES = [1, 2, 31, 33, 4095, 4097, 8191, 8192, 8193, 16383, 16384, 16385, 3 * 16384 + 5, 100003]


def dev(q, a):
    return ca.DeviceArray.from_host(np.ascontiguousarray(a), q)


def same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


def both(dt, cmp, c, fn="id", a=0, negate=False):
    """The same predicate for the product and for the oracle."""
    return ca.predicate(dt, cmp, c, fn, a, negate), O.pred(dt, cmp, c, fn, a, negate)


def data(dt, n, rng=RNG):
    dt = np.dtype(dt)
    if dt.kind == "f":
        x = rng.integers(-50, 50, n).astype(dt) * np.array(0.5, dt)
        if n > 16:
            idx = rng.integers(0, n, max(2, n // 64))
            x[idx] = rng.choice(np.array([np.nan, -np.nan, 0.0, -0.0, np.inf, -np.inf], dt), idx.size)
        return x
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, n, endpoint=True, dtype=dt)


def preds_for(dt):
    dt = np.dtype(dt)
    if dt.kind == "f":
        return [("<", 0.0, "id", 0), (">", 3.0, "mul", 0.5), ("!=", 0, "signbit", 0), ("==", 0.0, "id", 0)]
    info = np.iinfo(dt)
    mid = int(info.min) // 2 + int(info.max) // 2
    return [(">", mid, "id", 0), ("==", 1, "mod", 2), ("!=", 0, "and", 6), ("<", 0, "mul", 3), ("!=", 0, "mod", 7)]


# ============================================================ golden vectors
def test_golden_selection(queue, golden):
    seen = set()
    for c in golden:
        algo = c["algo"]
        if algo not in ("copy_if", "copy_index_if", "unique", "unique_copy", "partition", "remove", "count_if"):
            continue
        seen.add(algo)
        dt = c["dtype"]
        x = np.array(c["input"], dtype=dt)
        p = c.get("pred")
        pr = ca.predicate(dt, p["cmp"], p["c"], p["fn"], p["a"], p["negate"]) if p else None
        if algo in ("copy_if", "copy_index_if"):
            idx = algo == "copy_index_if"
            # the reference fills the output with -1 / 42 first and checks the tail is untouched
            fill = np.full(x.size, 77, np.uint32 if idx else dt)
            out = dev(queue, fill)
            k = ca.copy_if(queue, dev(queue, x), out, pr, ca.SELECT_INDEX_U32 if idx else ca.SELECT_VALUES)
            got = out.to_host(queue)
            assert k == len(c["expected"]), c["id"]
            assert got[:k].tolist() == c["expected"] and np.all(got[k:] == 77), c["id"]
        elif algo == "count_if":
            assert ca.count_if(queue, dev(queue, x), pr) == c["expected"], c["id"]
        elif algo == "unique_copy":
            out = dev(queue, np.zeros(x.size, dt))
            k = ca.unique(queue, dev(queue, x), out)
            assert out.to_host(queue)[:k].tolist() == c["expected"], c["id"]
        elif algo == "unique":
            d = dev(queue, x)
            k = ca.unique(queue, d, d)  # in place
            got = d.to_host(queue)
            assert k == c["expected_count"], c["id"]
            if "expected" in c:
                assert same(got, np.array(c["expected"], dt)), c["id"]
            if "expected_prefix" in c:
                assert same(got[:k], np.array(c["expected_prefix"], dt)), c["id"]
        elif algo == "partition":
            d = dev(queue, x)
            k = ca.partition(queue, d, d, None, pr)  # stable_partition in place
            assert k == c["expected_count"], c["id"]
            assert same(d.to_host(queue), np.array(c["expected"], dt)), c["id"]
        elif algo == "remove":
            d = dev(queue, x)
            for s in c["steps"]:
                pn = ca.predicate(dt, "==", s["value"], negate=True)
                k = ca.copy_if(queue, d, d, pn, n=s["len"])  # remove() in place on the prefix
                assert k == s["count"] and d.to_host(queue)[:k].tolist() == s["expected"], (c["id"], s)
    assert seen == {"copy_if", "copy_index_if", "unique", "unique_copy", "partition", "remove", "count_if"}


# ========================================================== oracle parity
@pytest.mark.parametrize("dt", ALL)
def test_copy_if_matches_oracle(queue, dt):
    for n in SIZES:
        x = data(dt, n)
        d = dev(queue, x)
        for cmp, c, fn, a in preds_for(dt):
            for neg in (False, True):
                pc, po = both(dt, cmp, c, fn, a, neg)
                want = O.copy_if(x, po)
                out = dev(queue, np.zeros(n, dt))
                k = ca.copy_if(queue, d, out, pc)
                assert k == want.size, (n, cmp, c, fn, a, neg)
                assert same(out.to_host(queue)[:k], want), (n, cmp, c, fn, a, neg)
        # indices, both widths, for one predicate
        pc, po = both(dt, *preds_for(dt)[0][:2])
        for kind, idt in ((ca.SELECT_INDEX_U32, np.uint32), (ca.SELECT_INDEX_U64, np.uint64)):
            out = dev(queue, np.zeros(n, idt))
            k = ca.copy_if(queue, d, out, pc, kind)
            assert same(out.to_host(queue)[:k], O.copy_if(x, po, kind)), (n, kind)


@pytest.mark.parametrize("dt", ALL)
def test_count_if_matches_oracle(queue, dt):
    for n in SIZES + [1 << 20, (1 << 20) + 3]:
        x = data(dt, n)
        d = dev(queue, x)
        for cmp, c, fn, a in preds_for(dt):
            pc, po = both(dt, cmp, c, fn, a)
            assert ca.count_if(queue, d, pc) == O.count_if(x, po), (n, cmp, c, fn, a)


def runs(dt, n, mean, rng=RNG):
    """Values in runs of random length (mean `mean`), so unique has work."""
    dt = np.dtype(dt)
    starts = np.sort(rng.choice(n, size=max(1, n // mean), replace=False)) if n > 1 else np.array([0])
    ids = np.zeros(n, np.int64)
    ids[starts] = 1
    ids = np.cumsum(ids)
    vals = data(dt, ids.max() + 2, rng)
    return vals[ids]


@pytest.mark.parametrize("dt", ALL)
def test_unique_matches_oracle(queue, dt):
    for n in SIZES:
        for mean in (1, 3, 50, 5000):
            x = runs(dt, n, mean)
            want = O.unique_copy(x)
            out = dev(queue, np.zeros(n, dt))
            k = ca.unique(queue, dev(queue, x), out)
            assert k == want.size and same(out.to_host(queue)[:k], want), (n, mean)
            d = dev(queue, x)  # in place: the tail keeps the input
            k = ca.unique(queue, d, d)
            got = d.to_host(queue)
            assert same(got[:k], want) and same(got[k:], x[k:]), (n, mean, "in place")


def test_unique_in_place_runs_across_tiles(queue):
    # one early duplicate shifts every later output by one position, and runs
    # straddle the tile boundaries: the element before each tile is read as
    # input while earlier tiles write just below it
    n = 6 * 8192 + 11
    x = np.arange(n, dtype=np.int32)
    x[1] = x[0]
    for b in range(8192, n, 8192):  # tile and staging-round boundaries
        x[b] = x[b - 1]
        x[b + 1] = x[b - 1]
    want = O.unique_copy(x)
    d = dev(queue, x)
    k = ca.unique(queue, d, d)
    assert k == want.size and same(d.to_host(queue)[:k], want)


@pytest.mark.parametrize("dt", ALL)
def test_partition_matches_oracle(queue, dt):
    for n in SIZES:
        x = data(dt, n)
        for cmp, c, fn, a in preds_for(dt)[:3]:
            pc, po = both(dt, cmp, c, fn, a)
            wt, wf = O.partition_copy(x, po)
            ot, of = dev(queue, np.zeros(n, dt)), dev(queue, np.zeros(n, dt))
            k = ca.partition(queue, dev(queue, x), ot, of, pc)
            assert k == wt.size, (n, cmp)
            assert same(ot.to_host(queue)[:k], wt) and same(of.to_host(queue)[:n - k], wf), (n, cmp)
            ws, wk = O.stable_partition(x, po)
            d = dev(queue, x)
            k = ca.partition(queue, d, d, None, pc)
            assert k == wk and same(d.to_host(queue), ws), (n, cmp, "in place")
            o = dev(queue, np.zeros(n, dt))
            k = ca.partition(queue, dev(queue, x), o, None, pc)
            assert k == wk and same(o.to_host(queue), ws), (n, cmp, "out of place")


@pytest.mark.parametrize("sel", ["none", "all", "sparse", "dense"])
def test_selectivity_extremes(queue, sel):
    n = 5 * 8192 + 77
    x = RNG.integers(0, 1000, n).astype(np.int32)
    c = {"none": -1, "all": 1000, "sparse": 10, "dense": 990}[sel]
    pc, po = both(np.int32, "<", c)
    want = O.copy_if(x, po)
    out = dev(queue, np.full(n, -7, np.int32))
    k = ca.copy_if(queue, dev(queue, x), out, pc)
    got = out.to_host(queue)
    assert k == want.size and same(got[:k], want) and np.all(got[k:] == -7)
    ws, wk = O.stable_partition(x, po)
    d = dev(queue, x)
    assert ca.partition(queue, d, d, None, pc) == wk and same(d.to_host(queue), ws)


@pytest.mark.parametrize("dt", ["int8", "int16", "float32", "float64", "uint64"])
def test_unaligned_subranges(queue, dt):
    # buffer_iterator offsets: the range starts inside a 16-byte chunk
    n = 3 * 8192 + 29
    x = data(dt, n + 16)
    base = dev(queue, x)
    isz = np.dtype(dt).itemsize
    for off in (1, 3, 5, 15 // isz):
        sub = x[off:off + n]
        pc, po = both(dt, *preds_for(dt)[0][:2])
        out = dev(queue, np.zeros(n, dt))
        k = ca.copy_if(queue, base, out, pc, n=n, in_ptr=base.at(off))
        assert same(out.to_host(queue)[:k], O.copy_if(sub, po)), off
        assert ca.count_if(queue, base, pc, n=n, in_ptr=base.at(off)) == O.count_if(sub, po), off
        k = ca.unique(queue, base, out, n=n, in_ptr=base.at(off))
        assert same(out.to_host(queue)[:k], O.unique_copy(sub)), off


def test_empty_ranges_write_zero_counts(queue):
    x = dev(queue, np.arange(4, dtype=np.int32))
    pc = ca.predicate(np.int32, ">", 0)
    cnt = dev(queue, np.array([99], np.uint64))
    for call in (lambda: ca.copy_if(queue, x, x, pc, n=0, count=cnt),
                 lambda: ca.unique(queue, x, x, n=0, count=cnt),
                 lambda: ca.partition(queue, x, x, None, pc, n=0, count=cnt)):
        cnt = dev(queue, np.array([99], np.uint64))
        call()
        assert int(cnt.to_host(queue)[0]) == 0
    assert ca.count_if(queue, x, pc, n=0) == 0


def test_invalid_predicates_are_rejected(queue):
    x = dev(queue, np.arange(8, dtype=np.float32))
    xi = dev(queue, np.arange(8, dtype=np.int32))
    out = dev(queue, np.zeros(8, np.float32))
    with pytest.raises(ca.BcmiError):
        ca.copy_if(queue, x, out, ca.predicate(np.float32, "==", 0, "mod", 2))
    with pytest.raises(ca.BcmiError):
        ca.copy_if(queue, xi, xi, ca.predicate(np.int32, "==", 0, "mod", 0))
    with pytest.raises(ca.BcmiError):
        ca.count_if(queue, xi, ca.predicate(np.int32, "!=", 0, "signbit"))
    with pytest.raises(ca.BcmiError):  # indices over the input
        ca.copy_if(queue, xi, xi, ca.predicate(np.int32, ">", 0), ca.SELECT_INDEX_U32)


def test_copy_if_full_size_properties(queue):
    # the benchmark's size: 2^28 int32, half selected; checked against numpy
    n = 1 << 28
    x = O.gen_scan_i32(n)
    pc = ca.predicate(np.int32, ">=", 0)
    mask = x >= 0
    d = dev(queue, x)
    out = dev(queue, np.zeros(n, np.int32))
    k = ca.copy_if(queue, d, out, pc)
    assert k == int(mask.sum())
    assert np.array_equal(out.to_host(queue)[:k], x[mask])
    assert ca.count_if(queue, d, pc) == k
    del out
    o2 = dev(queue, np.zeros(n, np.int32))
    assert ca.partition(queue, d, o2, None, pc) == k
    got = o2.to_host(queue)
    assert np.array_equal(got[:k], x[mask]) and np.array_equal(got[k:], x[~mask])


def test_graph_replay_of_compaction_and_mapped_reduction(queue):
    # captured once, replayed with new input contents and eager work between
    # replays: every replay must advance the look-back epochs and match
    n = 3 * 16384 + 77
    pc, po = both(np.int32, ">=", 0)
    x = RNG.integers(-100, 100, n).astype(np.int32)
    d = dev(queue, x)
    out = dev(queue, np.zeros(n, np.int32))
    part = dev(queue, np.zeros(n, np.int32))
    uq = dev(queue, np.zeros(n, np.int32))
    red = dev(queue, np.zeros(1, np.int32))
    f64 = dev(queue, np.zeros(1, np.float64))
    c1, c2, c3 = (dev(queue, np.zeros(1, np.uint64)) for _ in range(3))
    mp = ca.mapping(np.int32, "abs", "minus")

    def work():
        ca.copy_if(queue, d, out, pc, count=c1)
        ca.partition(queue, d, part, None, pc, count=c2)
        ca.unique(queue, d, uq, count=c3)
        ca.transform_reduce(queue, np.int32, d, out, n, mp, red, ca.MAX, f64=f64)

    work()  # sizes the queue's scratch and look-back state before capture
    queue.finish()
    g = ca.Graph(queue)
    with g:
        work()
    for rep in range(3):
        x = (RNG.integers(-100, 100, n) // (rep + 1)).astype(np.int32)  # runs for unique
        ca.check(ca.lib().bcmi_memcpy(queue.h, d.ptr, x.ctypes.data_as(ca.C.c_void_p), x.nbytes, ca.H2D, 1))
        ca.copy_if(queue, d, uq, pc)  # eager work on the same queue between replays
        g.launch()
        queue.finish()
        want = O.copy_if(x, po)
        k = int(c1.to_host(queue)[0])
        got = out.to_host(queue)
        assert k == want.size and same(got[:k], want), rep
        ws, wk = O.stable_partition(x, po)
        assert int(c2.to_host(queue)[0]) == wk and same(part.to_host(queue), ws), rep
        wu = O.unique_copy(x)
        assert int(c3.to_host(queue)[0]) == wu.size and same(uq.to_host(queue)[:wu.size], wu), rep
        wr = O.map_reduce(np.int32, x, got, n, O.mapping(np.int32, "abs", "minus"), O.MAX)
        assert red.to_host(queue)[0] == wr, rep
