"""bcmi_sort_noalloc (the allocation-free sort) against bcmi_sort.

Both entry points are called through the C-ABI on the same inputs; the
outputs must be bit-identical, on two calls in a row with the second one
smaller (the shared scratch keeps the first call's contents).  A size whose
temporaries exceed the scratch is refused before any work is enqueued."""
import ctypes as C

import numpy as np
import pytest

import compute_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu


def dev(q, a):
    return ca.DeviceArray.from_host(np.ascontiguousarray(a), q)


def _both(q, api, dtype, keys, vals, vb, desc):
    """Sort copies with bcmi_sort and bcmi_sort_noalloc; return both results."""
    L = ca.lib()
    out = []
    for fn in (L.bcmi_sort, L.bcmi_sort_noalloc):
        dk = dev(q, keys)
        dv = dev(q, vals) if vals is not None else None
        ca.check(fn(q.h, api, dtype, dk.ptr, dv.ptr if dv is not None else None, vb, keys.size, desc))
        out.append((dk.to_host(q), dv.to_host(q) if dv is not None else None))
    q.finish()
    return out


@pytest.mark.parametrize("n1,n2", [(3_000_017, 70_001), (1 << 20, 4097)])
def test_noalloc_sort_keys_matches_sort(queue, n1, n2):
    for n in (n1, n2):  # second call smaller
        k = O.gen_sort_u32(n, seed=0xB0057001 + n)
        (a, _), (b, _) = _both(queue, ca.SORT, ca.U32, k, None, 0, 0)
        assert np.array_equal(a, b), n
        assert np.array_equal(b, np.sort(k)), n


def test_noalloc_sort_by_key_float_matches_sort(queue):
    for n, desc in ((2_500_003, 0), (33_333, 1)):
        k, v = O.gen_sort_by_key_f32(n)
        (ak, av), (bk, bv) = _both(queue, ca.SORT_BY_KEY, ca.F32, k, v, 4, desc)
        assert np.array_equal(ak.view(np.uint32), bk.view(np.uint32)), n
        assert np.array_equal(av, bv), n
        ek, ev = O.radix_order_sort(k, v, descending=bool(desc))
        assert np.array_equal(bk.view(np.uint32), ek.view(np.uint32)), n
        assert np.array_equal(bv, ev), n


def test_noalloc_sort_u64_keys_u32_values_matches_sort(queue):
    rng = np.random.default_rng(7)
    for n in (1_000_003, 5000):
        k = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
        k[rng.integers(0, n, n // 10)] = k[0]
        v = np.arange(n, dtype=np.uint32)
        (ak, av), (bk, bv) = _both(queue, ca.STABLE_SORT_BY_KEY, ca.U64, k, v, 4, 0)
        assert np.array_equal(ak, bk) and np.array_equal(av, bv), n


def test_noalloc_sort_small_and_insertion_sizes(queue):
    for n in (31, 2, 40_000):
        k = O.gen_sort_u32(n, seed=n)
        (a, _), (b, _) = _both(queue, ca.SORT, ca.U32, k, None, 0, 0)
        assert np.array_equal(a, b), n


def test_noalloc_sort_refuses_oversize(queue):
    n = 20_000_000  # 80 MB of key temporaries: beyond the 64 MiB scratch
    d = ca.DeviceArray(n, np.uint32, queue)
    r = ca.lib().bcmi_sort_noalloc(queue.h, ca.SORT, ca.U32, d.ptr, None, 0, n, 0)
    assert r == -61  # CL_INVALID_BUFFER_SIZE
    assert b"scratch" in ca.lib().bcmi_last_error()
    queue.finish()
