"""CPU oracle for the data-parallel primitive core -- TEST INFRASTRUCTURE.

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package: it is the checker, never the product.
The product library (``compute_amd``) neither links nor calls it.

Thin numpy/ctypes wrappers over ``oracle/build/liboracle.so`` (built from
``bcmi_oracle.c`` by ``oracle/Makefile``).  Each C function cites the
reference (Boost.Compute 0.5.0) file:line it restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BCMI_ORACLE_BUILD selects another build directory (the sanitizer build,
# oracle/Makefile target "sanitize")
_BUILD = os.environ.get("BCMI_ORACLE_BUILD", os.path.join(_HERE, "build"))
_LIB_PATH = os.path.join(_BUILD, "liboracle.so")
_lib = None

# dtype codes (shared with the product C-ABI)
I8, U8, I16, U16, I32, U32, I64, U64, F32, F64 = range(1, 11)
PLUS, MULTIPLIES, MIN, MAX = 0, 1, 2, 3

NP2CODE = {np.dtype(np.int8): I8, np.dtype(np.uint8): U8, np.dtype(np.int16): I16,
           np.dtype(np.uint16): U16, np.dtype(np.int32): I32, np.dtype(np.uint32): U32,
           np.dtype(np.int64): I64, np.dtype(np.uint64): U64, np.dtype(np.float32): F32,
           np.dtype(np.float64): F64}


def kind_of(dt):
    """0 unsigned, 1 signed, 2 floating point."""
    dt = np.dtype(dt)
    if dt.kind == "f":
        return 2
    if dt.kind == "i":
        return 1
    return 0


class OrcPredicate(C.Structure):
    """f(_1) cmp c, optionally negated (orc_predicate in bcmi_oracle.h)."""
    _fields_ = [("fn", C.c_int32), ("cmp", C.c_int32), ("negate", C.c_int32), ("reserved", C.c_int32),
                ("operand", C.c_uint64), ("value", C.c_uint64)]


FN = {"id": 0, "mul": 1, "mod": 2, "and": 3, "signbit": 4}
CMP = {"<": 0, "<=": 1, ">": 2, ">=": 3, "==": 4, "!=": 5}


def _cmp_dtype(dt, fn):
    dt = np.dtype(dt)
    if fn == "signbit" or (dt.kind in "iu" and dt.itemsize < 4):
        return np.dtype(np.int32)
    return dt


def pred(dtype, cmp, c, fn="id", a=0, negate=False):
    """Predicate over elements of dtype; a and c are stored as one value of
    the comparison type (C's integer promotion; int for signbit)."""
    ct = _cmp_dtype(dtype, fn)
    def bits(v):
        return int(np.frombuffer(np.array([v]).astype(ct).tobytes().ljust(8, b"\0"), np.uint64)[0])
    return OrcPredicate(FN[fn], CMP[cmp], int(bool(negate)), 0, bits(a), bits(c))


class OrcMap(C.Structure):
    """y = g(a op b) or g(a - c) (orc_map in bcmi_oracle.h)."""
    _fields_ = [("op", C.c_int32), ("fn", C.c_int32), ("offset", C.c_uint64)]


MAP_FN = {"id": 0, "abs": 1, "square": 2}
MAP_OP = {"plus": 0, "multiplies": 1, "minus": 4}


def mapping(dtype, fn="id", op="multiplies", c=0):
    raw = np.array([c]).astype(dtype).tobytes().ljust(8, b"\0")
    return OrcMap(MAP_OP[op], MAP_FN[fn], int.from_bytes(raw, "little"))


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, sz, i = C.c_void_p, C.c_size_t, C.c_int
        L.orc_mt32_fill.argtypes = [C.c_uint32, vp, sz]
        L.orc_mt64_fill.argtypes = [C.c_uint64, vp, sz]
        L.orc_gen_sort_u32.argtypes = [C.c_uint32, vp, sz]
        L.orc_gen_sort_by_key_f32.argtypes = [vp, vp, sz]
        L.orc_gen_scan_i32.argtypes = [vp, sz]
        L.orc_gen_saxpy_f32.argtypes = [vp, vp, sz]
        L.orc_gen_lognormal_f32.argtypes = [vp, sz]
        L.orc_gen_segments.argtypes = [vp, sz, sz]
        L.orc_gen_segments.restype = sz
        L.orc_gen_segment_data.argtypes = [vp, vp, sz]
        L.orc_gen_runs_i32_f32.argtypes = [vp, vp, sz, sz]
        L.orc_gen_runs_i32_f32.restype = sz
        L.orc_radix_key.argtypes = [C.c_uint64, i, i, i]
        L.orc_radix_key.restype = C.c_uint64
        for f in ("orc_radix_order_sort", "orc_insertion_sort", "orc_compare_stable_sort"):
            getattr(L, f).argtypes = [vp, vp, sz, i, i, i, i]
        L.orc_sort_dispatch.argtypes = [i, vp, vp, sz, i, i, i, i]
        L.orc_segmented_sort.argtypes = [vp, vp, vp, sz, i, i, i, i]
        L.orc_scan.argtypes = [i, i, vp, vp, sz, vp, i]
        L.orc_scan_f32_dd.argtypes = [vp, vp, sz, i, C.c_double]
        L.orc_segmented_scan_i32keys.argtypes = [vp, i, vp, vp, sz, i, vp]
        L.orc_segmented_scan_f32_dd.argtypes = [vp, vp, vp, sz, i]
        L.orc_reduce.argtypes = [i, i, vp, sz, vp]
        L.orc_sum_f32_dd.argtypes = [vp, sz, C.POINTER(C.c_double)]
        L.orc_sum_f32_dd.restype = C.c_double
        L.orc_sum_f64_dd.argtypes = [vp, sz, C.POINTER(C.c_double)]
        L.orc_sum_f64_dd.restype = C.c_double
        L.orc_reduce_gpu_order_f32.argtypes = [vp, sz]
        L.orc_reduce_gpu_order_f32.restype = C.c_float
        L.orc_reduce_cpu_order_f32.argtypes = [vp, sz, C.c_uint]
        L.orc_reduce_cpu_order_f32.restype = C.c_float
        L.orc_scan_cpu_order.argtypes = [i, vp, vp, sz, i, vp, C.c_uint]
        L.orc_scan_cpu_order.restype = i
        L.orc_accumulate_f32.argtypes = [vp, sz, C.c_float]
        L.orc_accumulate_f32.restype = C.c_float
        L.orc_scan_gpu_order_f32.argtypes = [vp, vp, sz, i]
        L.orc_reduce_by_key.argtypes = [i, i, i, vp, vp, sz, vp, vp]
        L.orc_reduce_by_key.restype = sz
        L.orc_reduce_by_key_f32_dd.argtypes = [vp, vp, sz, vp, vp]
        L.orc_reduce_by_key_f32_dd.restype = sz
        L.orc_saxpy_f32.argtypes = [C.c_float, vp, vp, vp, sz]
        L.orc_transform.argtypes = [i, i, vp, vp, vp, sz]
        pp = C.POINTER(OrcPredicate)
        L.orc_pred_eval.argtypes = [i, vp, sz, pp]
        L.orc_pred_eval.restype = i
        L.orc_copy_if.argtypes = [i, vp, sz, pp, i, vp]
        L.orc_copy_if.restype = sz
        L.orc_count_if.argtypes = [i, vp, sz, pp]
        L.orc_count_if.restype = sz
        L.orc_unique_copy.argtypes = [i, vp, sz, vp]
        L.orc_unique_copy.restype = sz
        L.orc_partition_copy.argtypes = [i, vp, sz, pp, vp, vp]
        L.orc_partition_copy.restype = sz
        L.orc_stable_partition.argtypes = [i, vp, sz, pp]
        L.orc_stable_partition.restype = sz
        pm = C.POINTER(OrcMap)
        L.orc_map_reduce.argtypes = [i, vp, C.c_int64, vp, C.c_int64, i, sz, pm, i, vp, vp]
        L.orc_map_sum_dd.argtypes = [i, vp, C.c_int64, vp, C.c_int64, i, sz, pm, C.POINTER(C.c_double)]
        L.orc_map_sum_dd.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ------------------------------------------------------------ generators
def mt19937(seed, n):
    out = np.empty(n, np.uint32)
    lib().orc_mt32_fill(seed, _p(out), n)
    return out


def mt19937_64(seed, n):
    out = np.empty(n, np.uint64)
    lib().orc_mt64_fill(seed, _p(out), n)
    return out


def gen_sort_u32(n, seed=0xB0057001):
    out = np.empty(n, np.uint32)
    lib().orc_gen_sort_u32(seed, _p(out), n)
    return out


def gen_sort_by_key_f32(n):
    k = np.empty(n, np.float32)
    v = np.empty(n, np.uint32)
    lib().orc_gen_sort_by_key_f32(_p(k), _p(v), n)
    return k, v


def gen_scan_i32(n):
    out = np.empty(n, np.int32)
    lib().orc_gen_scan_i32(_p(out), n)
    return out


def gen_saxpy_f32(n):
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    lib().orc_gen_saxpy_f32(_p(x), _p(y), n)
    return x, y


def gen_lognormal_f32(n):
    out = np.empty(n, np.float32)
    lib().orc_gen_lognormal_f32(_p(out), n)
    return out


def gen_segments(nseg, max_len=1 << 20):
    off = np.empty(nseg + 1, np.uint64)
    lib().orc_gen_segments(_p(off), nseg, max_len)
    return off


def gen_segment_data(n):
    k = np.empty(n, np.uint64)
    v = np.empty(n, np.uint32)
    lib().orc_gen_segment_data(_p(k), _p(v), n)
    return k, v


def gen_runs_i32_f32(n, max_run=4096):
    k = np.empty(n, np.int32)
    v = np.empty(n, np.float32)
    lib().orc_gen_runs_i32_f32(_p(k), _p(v), n, max_run)
    return k, v


# ----------------------------------------------------------------- sorts
def radix_key(bits, dtype, descending=False):
    dt = np.dtype(dtype)
    return lib().orc_radix_key(int(bits), dt.itemsize, kind_of(dt), int(descending))


def _sortcall(fn, keys, vals, descending, *lead):
    k = np.ascontiguousarray(keys).copy()
    v = None if vals is None else np.ascontiguousarray(vals).copy()
    fn(*lead, _p(k), _p(v), k.size, k.dtype.itemsize, kind_of(k.dtype), int(descending),
       0 if v is None else v.dtype.itemsize)
    return k if v is None else (k, v)


def radix_order_sort(keys, vals=None, descending=False):
    return _sortcall(lib().orc_radix_order_sort, keys, vals, descending)


def insertion_sort(keys, vals=None, descending=False):
    return _sortcall(lib().orc_insertion_sort, keys, vals, descending)


def compare_stable_sort(keys, vals=None, descending=False):
    return _sortcall(lib().orc_compare_stable_sort, keys, vals, descending)


SORT, SORT_BY_KEY, STABLE_SORT, STABLE_SORT_BY_KEY = 0, 1, 2, 3


def sort_dispatch(api, keys, vals=None, descending=False):
    return _sortcall(lib().orc_sort_dispatch, keys, vals, descending, api)


def segmented_sort(keys, vals, offsets, descending=False):
    k = keys.copy()
    v = None if vals is None else vals.copy()
    off = np.ascontiguousarray(offsets, np.uint64)
    lib().orc_segmented_sort(_p(k), _p(v), _p(off), off.size - 1, k.dtype.itemsize,
                             kind_of(k.dtype), int(descending),
                             0 if v is None else v.dtype.itemsize)
    return k if v is None else (k, v)


# ------------------------------------------------------------------ scans
def scan(x, exclusive, init=0, op=PLUS):
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    ini = np.array([init], dtype=x.dtype)
    lib().orc_scan(NP2CODE[x.dtype], op, _p(x), _p(out), x.size, _p(ini), int(exclusive))
    return out


def scan_f32_dd(x, exclusive, init=0.0):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.size, np.float64)
    lib().orc_scan_f32_dd(_p(x), _p(out), x.size, int(exclusive), float(init))
    return out


def segmented_scan(keys, vals, exclusive, init=0):
    keys = np.ascontiguousarray(keys, np.int32)
    vals = np.ascontiguousarray(vals)
    out = np.empty_like(vals)
    ini = np.array([init], dtype=vals.dtype)
    lib().orc_segmented_scan_i32keys(_p(keys), NP2CODE[vals.dtype], _p(vals), _p(out),
                                     vals.size, int(exclusive), _p(ini))
    return out


def segmented_scan_f32_dd(keys, vals, exclusive):
    keys = np.ascontiguousarray(keys, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    out = np.empty(vals.size, np.float64)
    lib().orc_segmented_scan_f32_dd(_p(keys), _p(vals), _p(out), vals.size, int(exclusive))
    return out


# --------------------------------------------------------------- reduce
def reduce(x, op=PLUS):
    x = np.ascontiguousarray(x)
    out = np.empty(1, x.dtype)
    lib().orc_reduce(NP2CODE[x.dtype], op, _p(x), x.size, _p(out))
    return out[0]


def sum_dd(x):
    """Double-double sum: returns (hi, lo)."""
    lo = C.c_double()
    if x.dtype == np.float32:
        hi = lib().orc_sum_f32_dd(_p(np.ascontiguousarray(x)), x.size, C.byref(lo))
    else:
        hi = lib().orc_sum_f64_dd(_p(np.ascontiguousarray(x, np.float64)), x.size, C.byref(lo))
    return hi, lo.value


def reduce_gpu_order_f32(x):
    return np.float32(lib().orc_reduce_gpu_order_f32(_p(np.ascontiguousarray(x, np.float32)), x.size))


def reduce_cpu_order_f32(x, cu):
    return np.float32(lib().orc_reduce_cpu_order_f32(_p(np.ascontiguousarray(x, np.float32)), x.size, cu))


def scan_cpu_order(x, exclusive, cu, init=0, fill=0):
    """scan_on_cpu's association (plus, float32 or int32) with cu compute
    units; elements the reference leaves unwritten keep `fill`.  Raises on
    inputs whose result the reference leaves undefined."""
    x = np.ascontiguousarray(x)
    assert x.dtype in (np.float32, np.int32)
    out = np.full_like(x, fill)
    ini = np.array([init], dtype=x.dtype)
    r = lib().orc_scan_cpu_order(NP2CODE[x.dtype], _p(x), _p(out), x.size, int(exclusive), _p(ini), cu)
    if r != 0:
        raise ValueError("scan_on_cpu result undefined for n=%d, cu=%d" % (x.size, cu))
    return out


def accumulate_f32(x, init=0.0):
    return np.float32(lib().orc_accumulate_f32(_p(np.ascontiguousarray(x, np.float32)), x.size, init))


def scan_gpu_order_f32(x, exclusive):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().orc_scan_gpu_order_f32(_p(x), _p(out), x.size, int(exclusive))
    return out


def reduce_by_key(keys, vals, op=PLUS):
    keys = np.ascontiguousarray(keys)
    vals = np.ascontiguousarray(vals)
    ko = np.empty_like(keys)
    vo = np.empty_like(vals)
    r = lib().orc_reduce_by_key(NP2CODE[keys.dtype], NP2CODE[vals.dtype], op, _p(keys), _p(vals),
                                keys.size, _p(ko), _p(vo))
    return ko[:r], vo[:r]


def reduce_by_key_f32_dd(keys, vals):
    keys = np.ascontiguousarray(keys, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    ko = np.empty_like(keys)
    so = np.empty(keys.size, np.float64)
    r = lib().orc_reduce_by_key_f32_dd(_p(keys), _p(vals), keys.size, _p(ko), _p(so))
    return ko[:r], so[:r]


def saxpy(a, x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    lib().orc_saxpy_f32(a, _p(x), _p(y), _p(out), x.size)
    return out


def transform(a, b, op):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, a.dtype)
    out = np.empty_like(a)
    lib().orc_transform(NP2CODE[a.dtype], op, _p(a), _p(b), _p(out), a.size)
    return out


# ------------------------------------------------------------- selection
def pred_eval(x, p):
    x = np.ascontiguousarray(x)
    return np.array([lib().orc_pred_eval(NP2CODE[x.dtype], _p(x), i, C.byref(p)) for i in range(x.size)], bool)


def copy_if(x, p, out_kind=0):
    x = np.ascontiguousarray(x)
    odt = {0: x.dtype, 1: np.uint32, 2: np.uint64}[out_kind]
    out = np.zeros(x.size, odt)
    k = lib().orc_copy_if(NP2CODE[x.dtype], _p(x), x.size, C.byref(p), out_kind, _p(out))
    return out[:k]


def count_if(x, p):
    x = np.ascontiguousarray(x)
    return int(lib().orc_count_if(NP2CODE[x.dtype], _p(x), x.size, C.byref(p)))


def unique_copy(x):
    x = np.ascontiguousarray(x)
    out = np.zeros(x.size, x.dtype)
    k = lib().orc_unique_copy(NP2CODE[x.dtype], _p(x), x.size, _p(out))
    return out[:k]


def partition_copy(x, p):
    x = np.ascontiguousarray(x)
    t = np.zeros(x.size, x.dtype)
    f = np.zeros(x.size, x.dtype)
    k = lib().orc_partition_copy(NP2CODE[x.dtype], _p(x), x.size, C.byref(p), _p(t), _p(f))
    return t[:k], f[:x.size - k]


def stable_partition(x, p):
    d = np.ascontiguousarray(x).copy()
    k = lib().orc_stable_partition(NP2CODE[d.dtype], _p(d), d.size, C.byref(p))
    return d, int(k)


# ------------------------------------------------- transform_reduce
def _seq(x):
    """(pointer, start) of an array or of a counting sequence ('count', start)."""
    if isinstance(x, tuple):
        return None, int(x[1])
    return _p(x), 0


def map_reduce(dtype, a, b, n, m, rop, init=None):
    dt = np.dtype(dtype)
    (pa, a0), (pb, b0) = _seq(a), (_seq(b) if b is not None else (None, 0))
    out = np.zeros(1, dt)
    ini = None
    if init is not None:
        ini_arr = np.array([init], dt)
        ini = _p(ini_arr)
    lib().orc_map_reduce(NP2CODE[dt], pa, a0, pb, b0, int(b is not None), n, C.byref(m), rop, ini, _p(out))
    return out[0]


def map_sum_dd(dtype, a, b, n, m):
    dt = np.dtype(dtype)
    (pa, a0), (pb, b0) = _seq(a), (_seq(b) if b is not None else (None, 0))
    lo = C.c_double()
    hi = lib().orc_map_sum_dd(NP2CODE[dt], pa, a0, pb, b0, int(b is not None), n, C.byref(m), C.byref(lo))
    return hi, lo.value


# ------------------------------------------------------------ comparator sorts
# sort / stable_sort / sort_by_key with a user comparator: the reference's
# merge_sort_on_gpu (algorithm/detail/merge_sort_on_gpu.hpp:467-571) and, at
# n < 32, serial_insertion_sort (algorithm/detail/insertion_sort.hpp:26-72).
# The comparators are OpenCL C functions of two records; the forms restated
# here are the ones the reference's tests write (test/test_merge_sort_gpu.cpp:
# 239-242, 275-278, 311-314, 345-348; test/test_sort.cpp:260, 297-300;
# test/test_stable_sort.cpp:53-56, 76-79): key(a) < key(b) or key(a) > key(b)
# with key = a field, abs(field), abs(field + field), or length().
KEY_FORMS = {"field": 0, "abs": 1, "abs_sum2": 2, "length": 3}


def _as_rows(records, record_bytes):
    """The records as an (n, record_bytes) uint8 array."""
    raw = np.ascontiguousarray(records).view(np.uint8).reshape(-1)
    return raw.reshape(-1, record_bytes)


def _field(rows, dt, off):
    return np.ascontiguousarray(rows[:, off:off + dt.itemsize]).view(dt).reshape(-1)


def _magnitude_u64(s):
    """abs() of a signed integer array as its unsigned magnitude (uint64)."""
    u = s.astype(np.int64).view(np.uint64) if s.dtype != np.int64 else s.view(np.uint64)
    return np.where(s < 0, ~u + np.uint64(1), u)


def comparator_keys(rows, dtype, form, off, second=0, count=0):
    """key(record) for every row, vectorised: a field; abs(field) (integer abs
    is the unsigned magnitude, abs(INT_MIN) = 2^31; fabs for floats);
    abs(field + second field) with C's integer promotion (8/16-bit integers
    add in int, 32/64-bit integers wrap); length() through the sum of squares
    in double, in component order."""
    dt = np.dtype(dtype)
    form = KEY_FORMS.get(form, form)
    x = _field(rows, dt, off)
    if form == 0:
        return x
    if form == 3:
        ss = np.zeros(x.size, np.float64)
        for k in range(count):
            c = _field(rows, dt, off + k * dt.itemsize).astype(np.float64)
            ss = ss + c * c
        return ss
    if form == 2:
        y = _field(rows, dt, second)
        if dt.kind == "f":
            x = x + y
        elif dt.itemsize < 4:
            x = x.astype(np.int64) + y.astype(np.int64)
        else:
            u = (x.view(dt.str.replace("i", "u")) + y.view(dt.str.replace("i", "u")))  # wraps
            x = u.view(dt)
    if dt.kind == "f":
        return np.abs(x)
    if x.dtype.kind == "u":
        return x.astype(np.uint64)
    return _magnitude_u64(x)


def _stable_order(key, greater):
    if not greater:
        return np.argsort(key, kind="stable")
    n = key.size
    return (n - 1 - np.argsort(key[::-1], kind="stable"))[::-1]


def comparator_sort(records, record_bytes, dtype, form="field", lhs=0, rhs=None, second=0, count=0,
                    greater=False, values=None):
    """The stable sort of the records by the comparator (the order every
    strict weak order gives; the reference's stable merge sort and insertion
    sort produce it, its unstable sort one of the orders that ties allow).
    A comparator reading different fields of its two operands (a.s0 < b.s3)
    is sorted pairwise by comparator_less_py.  Returns (records, values) as
    flat arrays of the input dtypes."""
    rows = _as_rows(records, record_bytes)
    rhs = lhs if rhs is None else rhs
    if rhs == lhs:
        order = _stable_order(comparator_keys(rows, dtype, form, lhs, second, count), greater)
    else:
        import functools
        less = comparator_less_py(dtype, form, lhs, rhs, second, count, greater)
        items = list(range(rows.shape[0]))

        def cmp(i, j):
            return -1 if less(rows[i], rows[j]) else (1 if less(rows[j], rows[i]) else 0)
        order = np.array(sorted(items, key=functools.cmp_to_key(cmp)), np.int64)
    out = np.ascontiguousarray(rows[order]).view(np.asarray(records).dtype).reshape(-1)
    if values is None:
        return out, None
    v = np.asarray(values)
    vrows = v.view(np.uint8).reshape(rows.shape[0], -1)
    return out, np.ascontiguousarray(vrows[order]).view(v.dtype).reshape(-1)


def comparator_less_py(dtype, form, lhs, rhs=None, second=0, count=0, greater=False):
    """The comparator as a Python function of two records (uint8 rows), in
    exact integer / double arithmetic: the OpenCL C semantics spelled out
    independently of comparator_keys (Python ints, explicit wrap-around)."""
    import struct
    dt = np.dtype(dtype)
    form = KEY_FORMS.get(form, form)
    rhs = lhs if rhs is None else rhs
    code = {"i1": "b", "u1": "B", "i2": "h", "u2": "H", "i4": "i", "u4": "I", "i8": "q", "u8": "Q",
            "f4": "f", "f8": "d"}[dt.str[1:]]
    sz = dt.itemsize

    def fld(row, off):
        return struct.unpack_from("<" + code, bytes(row[off:off + sz]))[0]

    def wrap(v):
        if dt.kind == "f" or sz < 4:
            return v
        bits = 8 * sz
        v %= 1 << bits
        return v - (1 << bits) if dt.kind == "i" and v >= 1 << (bits - 1) else v

    def key(row, off):
        if form == 3:
            ss = 0.0
            for k in range(count):
                c = float(fld(row, off + k * sz))
                ss = ss + c * c
            return ss
        v = fld(row, off)
        if form == 0:
            return v
        if form == 2:
            if dt == np.float32:
                v = float(np.float32(v) + np.float32(fld(row, second)))
            else:
                v = wrap(v + fld(row, second))
        return abs(v)

    def less(a, b):
        ka, kb = key(a, lhs), key(b, rhs)
        return ka > kb if greater else ka < kb
    return less


def insertion_sort_rows(rows, less, values=None):
    """serial_insertion_sort / serial_insertion_sort_by_key
    (algorithm/detail/insertion_sort.hpp:26-72, :87-142): one work-item moves
    each element left past every element it is less than -- stable."""
    data = [r.copy() for r in rows]
    vals = None if values is None else list(values)
    for i in range(1, len(data)):
        v = data[i]
        vv = None if vals is None else vals[i]
        j = i
        while j > 0 and less(v, data[j - 1]):
            data[j] = data[j - 1]
            if vals is not None:
                vals[j] = vals[j - 1]
            j -= 1
        data[j] = v
        if vals is not None:
            vals[j] = vv
    return np.array(data).reshape(len(data), -1), vals


# ------------------------------------------------- CPU-column baseline port
# cpu_column.c: the reference's CPU-device algorithms on OpenMP threads, the
# benchmark's timed CPU baseline (never part of the library).
_cc = None
_CC_PATH = os.path.join(_BUILD, "libcpucol.so")


def cpucol():
    global _cc
    if _cc is None:
        if not os.path.exists(_CC_PATH):
            build()
        L = C.CDLL(_CC_PATH)
        vp, sz, i = C.c_void_p, C.c_size_t, C.c_int
        L.cc_sort_u32.argtypes = [vp, vp, sz, i, i]
        L.cc_sort_by_key_f32_u32.argtypes = [vp, vp, sz, i, i]
        L.cc_sort_by_key_u64_u32.argtypes = [vp, vp, sz, i, i]
        L.cc_segmented_sort_u64_u32.argtypes = [vp, vp, vp, sz, i]
        L.cc_scan_i32.argtypes = [vp, vp, sz, i, C.c_int32, i]
        L.cc_reduce_f32.argtypes = [vp, sz, i]
        L.cc_reduce_f32.restype = C.c_float
        L.cc_accumulate_f32.argtypes = [vp, sz, C.c_float]
        L.cc_accumulate_f32.restype = C.c_float
        L.cc_reduce_by_key_i32_f32.argtypes = [vp, vp, sz, vp, vp]
        L.cc_reduce_by_key_i32_f32.restype = sz
        L.cc_saxpy_f32.argtypes = [C.c_float, vp, vp, vp, sz, i]
        _cc = L
    return _cc


def cc_sort(keys, vals=None, threads=1):
    """merge_sort_on_cpu / merge_sort_by_key_on_cpu on `threads` threads
    (uint32 keys alone, float32 or uint64 keys with uint32 values)."""
    k = np.ascontiguousarray(keys).copy()
    if vals is None:
        assert k.dtype == np.uint32
        cpucol().cc_sort_u32(_p(k), None, k.size, threads, 1)
        return k
    v = np.ascontiguousarray(vals, np.uint32).copy()
    fn = {np.dtype(np.float32): cpucol().cc_sort_by_key_f32_u32,
          np.dtype(np.uint64): cpucol().cc_sort_by_key_u64_u32}[k.dtype]
    fn(_p(k), _p(v), k.size, threads, 0)
    return k, v


def cc_segmented_sort(keys, vals, offsets, threads=1):
    k, v = keys.copy(), vals.copy()
    off = np.ascontiguousarray(offsets, np.uint64)
    cpucol().cc_segmented_sort_u64_u32(_p(k), _p(v), _p(off), off.size - 1, threads)
    return k, v


def cc_scan_i32(x, exclusive, init=0, threads=1):
    x = np.ascontiguousarray(x, np.int32)
    out = np.empty_like(x)
    cpucol().cc_scan_i32(_p(x), _p(out), x.size, int(exclusive), int(init), threads)
    return out


def cc_reduce_f32(x, threads=1):
    return np.float32(cpucol().cc_reduce_f32(_p(np.ascontiguousarray(x, np.float32)), x.size, threads))


def cc_accumulate_f32(x, init=0.0):
    return np.float32(cpucol().cc_accumulate_f32(_p(np.ascontiguousarray(x, np.float32)), x.size, init))


def cc_reduce_by_key(keys, vals):
    keys = np.ascontiguousarray(keys, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    ko, vo = np.empty_like(keys), np.empty_like(vals)
    r = cpucol().cc_reduce_by_key_i32_f32(_p(keys), _p(vals), keys.size, _p(ko), _p(vo))
    return ko[:r], vo[:r]


def cc_saxpy(a, x, y, threads=1):
    out = np.empty_like(x)
    cpucol().cc_saxpy_f32(a, _p(x), _p(y), _p(out), x.size, threads)
    return out
