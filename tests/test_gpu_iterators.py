"""Iterator inputs through the C-ABI (bcmi_scan_iter, bcmi_transform_iter):
counting / constant sequences, element maps (offset, abs, square,
comparisons) and OpenCL vector elements, against the oracle applied to the
same sequences written out on the host.  Integers bit-exact; compensated
float scans within the bound of the double-double oracle."""
import numpy as np
import pytest

import compute_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)
DT = {"int32": ca.I32, "uint32": ca.U32, "int64": ca.I64, "uint64": ca.U64, "float32": ca.F32,
      "float64": ca.F64}
OPS = {"plus": (ca.PLUS, O.PLUS), "min": (ca.MIN, O.MIN), "max": (ca.MAX, O.MAX),
       "multiplies": (ca.MULTIPLIES, O.MULTIPLIES)}


def dev(q, a):
    return ca.DeviceArray.from_host(np.ascontiguousarray(a), q)


def host_sequence(dt, n, kind, value, arr=None):
    dt = np.dtype(dt)
    if kind == ca.ITER_BUFFER:
        return arr.copy()
    if kind == ca.ITER_CONSTANT:
        return np.full(n, value, dt)
    with np.errstate(over="ignore"):
        return (np.array(value, dt) + np.arange(n).astype(dt)).astype(dt)


def host_map(x, fn, c):
    with np.errstate(over="ignore", invalid="ignore"):
        if fn >= ca.MAP_CMP:
            ops = [np.less, np.less_equal, np.greater, np.greater_equal, np.equal, np.not_equal]
            return ops[fn - ca.MAP_CMP](x, np.array(c, x.dtype)).astype(x.dtype)
        t = (x - np.array(c, x.dtype)).astype(x.dtype)
        if fn == ca.MAP_ABS:
            return np.abs(t).astype(x.dtype) if x.dtype.kind != "u" else t
        if fn == ca.MAP_SQUARE:
            return (t * t).astype(x.dtype)
        return t


@pytest.mark.parametrize("dt", ["int32", "uint32", "int64", "uint64"])
@pytest.mark.parametrize("kind", [ca.ITER_COUNTING, ca.ITER_CONSTANT, ca.ITER_BUFFER])
@pytest.mark.parametrize("fn,c", [(ca.MAP_ID, 0), (ca.MAP_ID, 7), (ca.MAP_ABS, 50), (ca.MAP_SQUARE, 3),
                                  (ca.MAP_CMP + 2, 20)])
def test_scan_iter_integers_exact(queue, dt, kind, fn, c):
    n = 70001
    if kind == ca.ITER_BUFFER and fn == ca.MAP_ID and c == 0:
        pytest.skip("plain arrays are bcmi_scan's cases")
    start = -30 if np.dtype(dt).kind == "i" else 5
    arr = RNG.integers(-100 if np.dtype(dt).kind == "i" else 0, 100, n).astype(dt)
    d_arr = dev(queue, arr)
    x = host_map(host_sequence(dt, n, kind, start, arr), fn, c)
    out = ca.DeviceArray(n, np.dtype(dt), queue)
    it = ca.iterator(dt, d_arr if kind == ca.ITER_BUFFER else None, kind, start, fn, c)
    for name in ("plus", "max", "min"):
        cop, oop = OPS[name]
        for excl in (False, True):
            ini = np.array([3], dt)
            ca.check(ca.lib().bcmi_scan_iter(queue.h, DT[dt], 1, cop, it, out.ptr, n,
                                             ini.ctypes.data_as(ca.C.c_void_p) if excl else None, int(excl), 0))
            exp = O.scan(x, excl, 3 if excl else 0, oop)
            assert np.array_equal(out.to_host(queue), exp), (name, excl)


@pytest.mark.parametrize("fn,c", [(ca.MAP_ID, 0.0), (ca.MAP_SQUARE, 0.0), (ca.MAP_SQUARE, 0.5), (ca.MAP_ABS, 1.0)])
def test_scan_iter_float_compensated(queue, fn, c):
    n = 200003
    arr = O.gen_lognormal_f32(n)
    d_arr = dev(queue, arr)
    x = host_map(arr, fn, np.float32(c))
    out = ca.DeviceArray(n, np.float32, queue)
    it = ca.iterator("float32", d_arr, ca.ITER_BUFFER, 0, fn, c)
    ca.check(ca.lib().bcmi_scan_iter(queue.h, ca.F32, 1, ca.PLUS, it, out.ptr, n, None, 0, 0))
    dd = O.scan_f32_dd(x, False)
    err = np.abs(out.to_host(queue).astype(np.float64) - dd)
    bound = np.spacing(np.abs(dd).astype(np.float32)).astype(np.float64) + \
        2.0 ** -50 * np.cumsum(np.abs(x.astype(np.float64)))
    assert np.all(err <= bound)


def test_scan_iter_counting_float64_exact(queue):
    n = 100000
    it = ca.iterator("float64", None, ca.ITER_COUNTING, 0.5)
    out = ca.DeviceArray(n, np.float64, queue)
    ca.check(ca.lib().bcmi_scan_iter(queue.h, ca.F64, 1, ca.PLUS, it, out.ptr, n, None, 0, ca.NAIVE))
    i = np.arange(1, n + 1, dtype=np.float64)
    assert np.array_equal(out.to_host(queue), i * (i - 1) / 2 + 0.5 * i)  # all exact in double


@pytest.mark.parametrize("dt,w", [("int32", 2), ("int32", 4), ("uint32", 2), ("float32", 4)])
@pytest.mark.parametrize("n", [5, 4097, 100003])
def test_scan_iter_vector_elements(queue, dt, w, n):
    if np.dtype(dt).kind == "f":
        a = RNG.integers(-1000, 1000, (n, w)).astype(dt)  # exact sums
    else:
        a = RNG.integers(0, 2 ** 31, (n, w)).astype(dt)
    d = ca.DeviceArray.from_host(a.reshape(-1), queue)
    out = ca.DeviceArray(n * w, np.dtype(dt), queue)
    it = ca.iterator(dt, d, ca.ITER_BUFFER)
    for excl in (False, True):
        ini = np.arange(1, w + 1).astype(dt)
        ca.check(ca.lib().bcmi_scan_iter(queue.h, DT[dt], w, ca.PLUS, it, out.ptr, n,
                                         ini.ctypes.data_as(ca.C.c_void_p) if excl else None, int(excl), 0))
        got = out.to_host(queue).reshape(n, w)
        for k in range(w):
            exp = O.scan(np.ascontiguousarray(a[:, k]), excl, ini[k] if excl else 0)
            assert np.array_equal(got[:, k], exp), (excl, k)


def test_scan_iter_rejects_bad_descriptors(queue):
    out = ca.DeviceArray(10, np.int32, queue)
    bad = ca.iterator("int32", None, ca.ITER_COUNTING, 0, fn=7)
    assert ca.lib().bcmi_scan_iter(queue.h, ca.I32, 1, ca.PLUS, bad, out.ptr, 10, None, 0, 0) == -30
    vec_counting = ca.iterator("int32", None, ca.ITER_COUNTING, 0)
    assert ca.lib().bcmi_scan_iter(queue.h, ca.I32, 2, ca.PLUS, vec_counting, out.ptr, 5, None, 0, 0) == -30
    assert ca.lib().bcmi_scan_iter(queue.h, ca.I8, 1, ca.PLUS, vec_counting, out.ptr, 5, None, 0, 0) == -30


@pytest.mark.parametrize("dt", ["int32", "int64", "float32", "float64"])
@pytest.mark.parametrize("op", ["plus", "minus", "multiplies", "min", "max"])
def test_transform_iter_binary(queue, dt, op):
    n = 50003
    arr = (RNG.normal(size=n) * 100).astype(dt)
    d_arr = dev(queue, arr)
    code = {"plus": ca.PLUS, "minus": ca.MINUS, "multiplies": ca.MULTIPLIES, "min": ca.MIN, "max": ca.MAX}[op]
    ocode = {"plus": O.PLUS, "minus": 4, "multiplies": O.MULTIPLIES, "min": O.MIN, "max": O.MAX}[op]
    a = ca.iterator(dt, None, ca.ITER_COUNTING, 3)
    b = ca.iterator(dt, d_arr, ca.ITER_BUFFER, 0, ca.MAP_ABS, 0)
    out = ca.DeviceArray(n, np.dtype(dt), queue)
    ca.check(ca.lib().bcmi_transform_iter(queue.h, DT[dt], 1, code, a, b, out.ptr, n))
    ha = host_sequence(dt, n, ca.ITER_COUNTING, 3)
    hb = host_map(arr, ca.MAP_ABS, 0)
    if op == "minus":  # the oracle's transform has the scan/reduce functors only
        with np.errstate(over="ignore"):
            exp = (ha - hb).astype(dt)
    else:
        exp = O.transform(ha, hb, ocode)
    assert np.array_equal(out.to_host(queue).view(np.uint8), exp.view(np.uint8))


def test_transform_iter_copy_and_length(queue):
    n = 1000
    out = ca.DeviceArray(n, np.int32, queue)
    c = ca.iterator("int32", None, ca.ITER_CONSTANT, 42)
    ca.check(ca.lib().bcmi_transform_iter(queue.h, ca.I32, 1, ca.PLUS, c, None, out.ptr, n))
    assert np.array_equal(out.to_host(queue), np.full(n, 42, np.int32))
    v = RNG.normal(size=(n, 4)).astype(np.float32)
    dv = ca.DeviceArray.from_host(v.reshape(-1), queue)
    lo = ca.DeviceArray(n, np.float32, queue)
    it = ca.iterator("float32", dv, ca.ITER_BUFFER, 0, ca.MAP_LENGTH)
    ca.check(ca.lib().bcmi_transform_iter(queue.h, ca.F32, 4, ca.PLUS, it, None, lo.ptr, n))
    exp = np.sqrt((v.astype(np.float64) ** 2).sum(axis=1))
    assert np.allclose(lo.to_host(queue), exp, rtol=3e-7, atol=0)
