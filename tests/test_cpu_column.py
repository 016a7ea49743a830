"""The benchmark's CPU baseline (oracle/cpu_column.c: the reference's
CPU-device algorithms run on OpenMP threads) against the single-threaded
oracle restatements it must agree with, bit for bit, at several thread
counts.  A baseline that computed something else would time the wrong
work."""
import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(31)


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("n", [1, 2, 511, 512, 513, 4097, 100003])
def test_cc_sort_u32(threads, n):
    k = RNG.integers(0, 2 ** 32, n, dtype=np.uint32)
    k[RNG.integers(0, n, n // 4)] = k[0]
    assert np.array_equal(O.cc_sort(k, threads=threads), np.sort(k))


def test_cc_sort_u32_merge_path_passes():
    """n >= 2^21: the last passes (<= 8 blocks) merge by merge path."""
    n = (1 << 21) + 777
    k = RNG.integers(0, 1 << 20, n, dtype=np.uint32)
    assert np.array_equal(O.cc_sort(k, threads=8), np.sort(k))


@pytest.mark.parametrize("threads", [1, 5])
@pytest.mark.parametrize("n", [3, 600, 65537])
def test_cc_sort_by_key_f32_is_the_cpu_column(threads, n):
    """operator< on floats: -0.0 == +0.0 keep input order (stable), which
    the oracle's compare_stable_sort restates."""
    k, v = O.gen_sort_by_key_f32(n)
    k = k[RNG.permutation(n)]
    k[RNG.integers(0, n, max(1, n // 10))] = np.float32(-0.0)
    k[RNG.integers(0, n, max(1, n // 10))] = np.float32(0.0)
    gk, gv = O.cc_sort(k, v, threads=threads)
    ek, ev = O.compare_stable_sort(k, v)
    assert np.array_equal(gk.view(np.uint32), ek.view(np.uint32))
    assert np.array_equal(gv, ev)


def test_cc_segmented_sort_u64():
    off = np.array([0, 1, 700, 700, 5000, 5003, 70000], np.uint64)
    k, v = O.gen_segment_data(int(off[-1]))
    k[100:200] = k[100]
    gk, gv = O.cc_segmented_sort(k, v, off, threads=4)
    ek, ev = O.segmented_sort(k, v, off)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


@pytest.mark.parametrize("threads", [1, 2, 7, 16])
@pytest.mark.parametrize("n", [5, 70000, 1 << 20])
def test_cc_scan_i32(threads, n):
    x = O.gen_scan_i32(n)
    for excl, init in ((True, 0), (True, -9), (False, 0)):
        got = O.cc_scan_i32(x, excl, init, threads=threads)
        assert np.array_equal(got, O.scan(x, excl, init)), (excl, init)


@pytest.mark.parametrize("threads", [1, 3, 16])
def test_cc_reduce_f32_is_reduce_on_cpu_order(threads):
    x = O.gen_lognormal_f32(300001)
    got = O.cc_reduce_f32(x, threads=threads)
    exp = O.reduce_cpu_order_f32(x, threads)
    assert got.view(np.uint32) == exp.view(np.uint32)


def test_cc_accumulate_and_reduce_by_key_and_saxpy():
    x = O.gen_lognormal_f32(100000)
    assert O.cc_accumulate_f32(x, 1.5).view(np.uint32) == O.accumulate_f32(x, 1.5).view(np.uint32)
    k, v = O.gen_runs_i32_f32(200000)
    gk, gv = O.cc_reduce_by_key(k, v)
    ek, ev = O.reduce_by_key(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv.view(np.uint32), ev.view(np.uint32))
    a, b = O.gen_saxpy_f32(100003)
    assert np.array_equal(O.cc_saxpy(2.5, a, b, threads=6).view(np.uint32), O.saxpy(2.5, a, b).view(np.uint32))
