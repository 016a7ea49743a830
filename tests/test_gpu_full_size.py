"""Every benchmark workload at its full size, checked on the GPU.

Sizes too large for the oracle to replay in full are checked through
properties that pin the result exactly:
  * sort_by_key of 2^27 float32 keys with payload = original index: the
    stable sort by radix(key) is the unique output in which the payload is
    a permutation, keys_out == keys_in[payload] bit for bit, the radix keys
    are non-decreasing and, among equal radix keys, the payload increases;
  * saxpy is compared bit for bit with the oracle (fmaf per element);
  * float scans / reduce_by_key / segmented scan are compared with the
    double-double oracle within the bound stated in each test.
"""
import numpy as np
import pytest

import compute_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu


def dev(q, a):
    return ca.DeviceArray.from_host(np.ascontiguousarray(a), q)


def radix_f32_asc(u):
    """radix() of float32 bit patterns, ascending (radix_sort.hpp:102-104)."""
    u = u.view(np.uint32)
    return u ^ ((np.uint32(0) - (u >> np.uint32(31))) | np.uint32(0x80000000))


def test_radix_f32_matches_oracle_key():
    probe = np.array([0.0, -0.0, 1.5, -1.5, np.inf, -np.inf, np.nan, 1e-45, -1e-45], np.float32)
    got = radix_f32_asc(probe.view(np.uint32))
    exp = [O.radix_key(int(b), np.float32) for b in probe.view(np.uint32)]
    assert [int(g) for g in got] == exp


def test_sort_by_key_f32_u32_2p27_stable(queue):
    """sort_by_key of 2^27 float32 keys in [-1e6, 1e6] with 1% signed zeros
    and duplicates, payload = index (the 68 B/pair benchmark workload, on
    the tuned 2^27 route)."""
    n = 1 << 27
    k, v = O.gen_sort_by_key_f32(n)
    assert np.array_equal(v, np.arange(n, dtype=np.uint32))
    dk, dv = dev(queue, k), dev(queue, v)
    ca.sort(queue, dk, dv, api=ca.SORT_BY_KEY)
    gk, gv = dk.to_host(queue), dv.to_host(queue)
    seen = np.zeros(n, bool)
    seen[gv] = True
    assert seen.all(), "payload is not a permutation"
    assert np.array_equal(gk.view(np.uint32), k.view(np.uint32)[gv]), "keys do not follow the payload"
    rk = radix_f32_asc(gk.view(np.uint32))
    assert np.all(rk[1:] >= rk[:-1]), "radix keys out of order"
    eq = rk[1:] == rk[:-1]
    assert np.all(gv[1:][eq] > gv[:-1][eq]), "equal keys not in input order"
    # signed zeros keep their bits and order: every -0.0 precedes every +0.0
    z = gk == 0
    sb = gk.view(np.uint32)[z] >> np.uint32(31)
    assert np.all(sb[1:] <= sb[:-1])


def test_exclusive_scan_i32_2p28_init(queue):
    """exclusive_scan of 2^28 int32 with a non-zero init (wrapping int32)."""
    n = 1 << 28
    x = O.gen_scan_i32(n)
    d = dev(queue, x)
    ca.scan(queue, d, d, True, init=-7)  # in place
    got = d.to_host(queue)
    exp = np.empty(n, np.int32)
    exp[0] = -7
    np.cumsum(x[:-1], dtype=np.int32, out=exp[1:])
    exp[1:] += np.int32(-7)
    assert np.array_equal(got, exp)


def test_saxpy_2p28_bit_exact(queue):
    """y = 2.5f*x + y on 2^28 float32 N(0,1): one fmaf per element."""
    n = 1 << 28
    x, y = O.gen_saxpy_f32(n)
    dx, dy = dev(queue, x), dev(queue, y)
    ca.saxpy(queue, 2.5, dx, dy, dy)
    got = dy.to_host(queue)
    assert np.array_equal(got.view(np.uint32), O.saxpy(2.5, x, y).view(np.uint32))


def _scan_bound(x, dd):
    # compensated within a tile + FP64 carry: one ulp of the result plus the
    # FP64 carry's rounding, 2^-50 * sum|x| of the prefix
    return np.spacing(np.abs(dd).astype(np.float32)).astype(np.float64) + \
        2.0 ** -50 * np.cumsum(np.abs(x.astype(np.float64)))


def test_inclusive_scan_f32_2p28_compensated(queue):
    """inclusive_scan of the 2^28 lognormal(0, 3) floats with random signs:
    every prefix within one ulp of its double-double value plus the FP64
    carry's rounding."""
    n = 1 << 28
    x = O.gen_lognormal_f32(n)
    out = ca.DeviceArray(n, np.float32, queue)
    ca.scan(queue, dev(queue, x), out, False)
    dd = O.scan_f32_dd(x, False)
    err = np.abs(out.to_host(queue).astype(np.float64) - dd)
    assert np.all(err <= _scan_bound(x, dd))


def test_accumulate_f32_2p28_with_init(queue):
    """accumulate(first, last, 1.5f) on the 2^28 lognormal floats: within 2
    ulp of the double-double value of init + sum."""
    n = 1 << 28
    x = O.gen_lognormal_f32(n)
    out = ca.DeviceArray(1, np.float32, queue)
    ca.reduce(queue, dev(queue, x), out, init=np.float32(1.5))
    hi, lo = O.sum_dd(np.r_[np.float32(1.5), x])
    ulp = float(np.spacing(np.float32(abs(hi))))
    assert abs(float(out.to_host(queue)[0]) - (hi + lo)) <= 2 * ulp


@pytest.fixture(scope="module")
def runs_2p27():
    return O.gen_runs_i32_f32(1 << 27)


def test_reduce_by_key_2p27(queue, runs_2p27):
    """reduce_by_key(plus) over 2^27 sorted int32 run keys (geometric runs,
    mean 32, max 4096) with float32 values: keys bit-exact, each run sum
    within one ulp plus the FP64 carry's rounding of its double-double sum."""
    keys, vals = runs_2p27
    n = keys.size
    ko = ca.DeviceArray(n, np.int32, queue)
    vo = ca.DeviceArray(n, np.float32, queue)
    r = ca.reduce_by_key(queue, dev(queue, keys), dev(queue, vals), ko, vo)
    ek, es = O.reduce_by_key_f32_dd(keys, vals)
    assert r == ek.size
    assert np.array_equal(ko.to_host(queue)[:r], ek)
    got = vo.to_host(queue)[:r].astype(np.float64)
    heads = np.r_[0, np.flatnonzero(np.diff(keys)) + 1]
    absum = np.add.reduceat(np.abs(vals.astype(np.float64)), heads)
    bound = np.spacing(np.abs(es).astype(np.float32)).astype(np.float64) + 2.0 ** -50 * absum
    assert np.all(np.abs(got - es) <= bound)


def test_segmented_exclusive_scan_2p27(queue, runs_2p27):
    """Segmented exclusive scan of the same 2^27 runs: every element within
    one ulp plus the FP64 carry's rounding of the double-double prefix of
    its run (the run restarts at 0)."""
    keys, vals = runs_2p27
    n = keys.size
    out = ca.DeviceArray(n, np.float32, queue)
    ca.segmented_scan(queue, dev(queue, keys), dev(queue, vals), out, True, init=0.0)
    dd = O.segmented_scan_f32_dd(keys, vals, True)
    err = np.abs(out.to_host(queue).astype(np.float64) - dd)
    bound = np.spacing(np.abs(dd).astype(np.float32)).astype(np.float64) + \
        2.0 ** -50 * np.cumsum(np.abs(vals.astype(np.float64)))
    assert np.all(err <= bound)
    heads = np.r_[0, np.flatnonzero(np.diff(keys)) + 1]
    assert np.all(out.to_host(queue)[heads] == 0.0)
