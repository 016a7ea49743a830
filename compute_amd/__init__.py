"""compute_amd -- MI355X-native data-parallel primitive core (Boost.Compute's
sort / scan / reduce / reduce_by_key / transform hot path).

The product is the C-ABI library ``libbcmi.so`` (declared in
``include/bcmi.h``) with hand-written gfx950 HIP kernels, and the header-only
C++ templates under ``include/boost/compute/`` that keep the reference's
signatures.  This module is a thin ctypes binding of that C-ABI for Python
callers, tests and the benchmark.  It never falls back to a CPU path: if the
library is missing or no gfx950 device is visible, calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbcmi.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

# enums from include/bcmi.h
I8, U8, I16, U16, I32, U32, I64, U64, F32, F64 = range(1, 11)
PLUS, MULTIPLIES, MIN, MAX, MINUS, DIVIDES = range(6)
SORT, SORT_BY_KEY, STABLE_SORT, STABLE_SORT_BY_KEY, RADIX_SORT = range(5)
COMPENSATED, NAIVE = 0, 1
H2D, D2H, D2D = 1, 2, 3
OUT_OF_RESOURCES, INVALID_VALUE, INVALID_BUFFER_SIZE = -5, -30, -61

DTYPE = {np.dtype(np.int8): I8, np.dtype(np.uint8): U8, np.dtype(np.int16): I16,
         np.dtype(np.uint16): U16, np.dtype(np.int32): I32, np.dtype(np.uint32): U32,
         np.dtype(np.int64): I64, np.dtype(np.uint64): U64, np.dtype(np.float32): F32,
         np.dtype(np.float64): F64}

EXPORTS = [
    "bcmi_last_error", "bcmi_get_version", "bcmi_device_count", "bcmi_device_get_info",
    "bcmi_queue_create", "bcmi_queue_destroy", "bcmi_queue_finish", "bcmi_queue_device",
    "bcmi_queue_native", "bcmi_queue_reserve", "bcmi_malloc", "bcmi_free", "bcmi_memcpy",
    "bcmi_fill", "bcmi_event_create", "bcmi_event_destroy", "bcmi_event_record",
    "bcmi_event_synchronize", "bcmi_event_elapsed_ms", "bcmi_sort", "bcmi_sort_noalloc", "bcmi_scan", "bcmi_reduce",
    "bcmi_reduce_by_key", "bcmi_transform", "bcmi_saxpy", "bcmi_segmented_sort",
    "bcmi_segmented_plan_create", "bcmi_segmented_sort_plan", "bcmi_segmented_plan_destroy",
    "bcmi_segmented_scan", "bcmi_graph_begin", "bcmi_graph_end", "bcmi_graph_launch",
    "bcmi_graph_destroy", "bcmi_tune_get", "bcmi_tune_set", "bcmi_tuning_cache_load",
    "bcmi_tuning_cache_save", "bcmi_profile_enable", "bcmi_profile_query", "bcmi_profile_reset",
    "bcmi_copy_if", "bcmi_count_if", "bcmi_unique", "bcmi_partition", "bcmi_transform_reduce",
    "bcmi_extract_keys", "bcmi_host_malloc", "bcmi_host_free", "bcmi_merge_sort", "bcmi_scan_iter",
    "bcmi_transform_iter", "bcmi_convert", "bcmi_debug_fault_probe",
]

SELECT_VALUES, SELECT_INDEX_U32, SELECT_INDEX_U64 = 0, 1, 2
FN = {"id": 0, "mul": 1, "mod": 2, "and": 3, "signbit": 4}
CMP = {"<": 0, "<=": 1, ">": 2, ">=": 3, "==": 4, "!=": 5}


MAP_FN = {"id": 0, "abs": 1, "square": 2}
MAP_OP = {"plus": PLUS, "minus": MINUS, "multiplies": MULTIPLIES}


class Seq(C.Structure):
    """bcmi_seq: a device array, or a counting sequence start + i (ptr NULL)."""
    _fields_ = [("ptr", C.c_void_p), ("start", C.c_int64)]


class Map(C.Structure):
    """bcmi_map: y = g(a op b) (binary) or g(a - c) (unary)."""
    _fields_ = [("op", C.c_int32), ("fn", C.c_int32), ("offset", C.c_uint64)]


ITER_BUFFER, ITER_COUNTING, ITER_CONSTANT = 0, 1, 2
MAP_ID, MAP_ABS, MAP_SQUARE, MAP_LENGTH, MAP_CMP = 0, 1, 2, 3, 8


class Iter(C.Structure):
    """bcmi_iter: an array, counting or constant sequence and its element map."""
    _fields_ = [("ptr", C.c_void_p), ("kind", C.c_int32), ("fn", C.c_int32), ("value", C.c_uint64),
                ("offset", C.c_uint64)]


def _elem_bits(dtype, v):
    a = np.zeros(1, np.uint64)
    a.view(np.uint8)[:np.dtype(dtype).itemsize] = np.array([v], dtype=dtype).view(np.uint8)
    return int(a[0])


def iterator(dtype, array=None, kind=None, value=0, fn=MAP_ID, offset=0):
    """Build a bcmi_iter: an array (DeviceArray or device pointer), or a
    counting / constant sequence starting at / equal to value."""
    if kind is None:
        kind = ITER_BUFFER if array is not None else ITER_COUNTING
    ptr = array.ptr if hasattr(array, "ptr") else array
    return Iter(ptr, kind, fn, _elem_bits(dtype, value), _elem_bits(dtype, offset))


def mapping(dtype, fn="id", op="multiplies", c=0):
    raw = np.array([c]).astype(np.dtype(dtype)).tobytes()
    return Map(MAP_OP[op], MAP_FN[fn], int.from_bytes(raw.ljust(8, b"\0"), "little"))


KEY_FORM = {"field": 0, "abs": 1, "abs_sum2": 2, "length": 3}


class Comparator(C.Structure):
    """bcmi_comparator: key(a at lhs_offset) < (or >) key(b at rhs_offset)."""
    _fields_ = [("dtype", C.c_int32), ("form", C.c_int32), ("lhs_offset", C.c_uint32),
                ("rhs_offset", C.c_uint32), ("offset2", C.c_uint32), ("count", C.c_uint32),
                ("greater", C.c_int32)]


def comparator(dtype, form="field", lhs=0, rhs=None, second=0, count=0, greater=False):
    """A comparator over fields of scalar type `dtype` at byte offsets of each
    record: form field | abs | abs_sum2 (abs(field + field at `second`)) |
    length (of `count` components)."""
    return Comparator(DTYPE[np.dtype(dtype)], KEY_FORM[form], lhs, lhs if rhs is None else rhs, second, count,
                      int(bool(greater)))


def counting(start):
    """A counting sequence (make_counting_iterator(start))."""
    return Seq(None, int(start))


class Predicate(C.Structure):
    """bcmi_predicate: f(_1) cmp c, optionally negated (not1)."""
    _fields_ = [("fn", C.c_int32), ("cmp", C.c_int32), ("negate", C.c_int32), ("reserved", C.c_int32),
                ("operand", C.c_uint64), ("value", C.c_uint64)]


def predicate(dtype, cmp, c, fn="id", a=0, negate=False):
    """Build a predicate over elements of `dtype`.  a and c are converted to
    the comparison type: int32 for 8/16-bit integers and for signbit (C's
    integer promotion / OpenCL's int result), else the element type."""
    dt = np.dtype(dtype)
    ct = np.dtype(np.int32) if (fn == "signbit" or (dt.kind in "iu" and dt.itemsize < 4)) else dt
    def bits(v):
        raw = np.array([v]).astype(ct).tobytes()
        return int.from_bytes(raw.ljust(8, b"\0"), "little")
    return Predicate(FN[fn], CMP[cmp], int(bool(negate)), 0, bits(a), bits(c))


class BcmiError(RuntimeError):
    """Mirrors boost::compute::opencl_error: carries the negative status."""

    def __init__(self, code, msg):
        super().__init__("bcmi error %d: %s" % (code, msg))
        self.code = code


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int),
                ("wavefront_size", C.c_int), ("clock_khz", C.c_int),
                ("global_mem_bytes", C.c_size_t), ("lds_bytes_per_block", C.c_size_t),
                ("pci_bus_id", C.c_int)]


_lib = None


def lib():
    """Load libbcmi.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BcmiError(-9999, "libbcmi.so not built: run __graft_entry__.build() or "
                                   "make -C compute_amd/csrc")
        L = C.CDLL(LIB_PATH)
        vp, sz, i, u64p = C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_uint64)
        L.bcmi_last_error.restype = C.c_char_p
        L.bcmi_get_version.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.bcmi_device_count.argtypes = [C.POINTER(C.c_int)]
        L.bcmi_device_get_info.argtypes = [i, C.POINTER(DeviceInfo)]
        L.bcmi_queue_create.argtypes = [i, C.POINTER(vp)]
        L.bcmi_queue_destroy.argtypes = [vp]
        L.bcmi_queue_finish.argtypes = [vp]
        L.bcmi_queue_device.argtypes = [vp, C.POINTER(C.c_int)]
        L.bcmi_queue_native.argtypes = [vp]
        L.bcmi_queue_native.restype = vp
        L.bcmi_queue_reserve.argtypes = [vp, sz, sz]
        L.bcmi_malloc.argtypes = [sz, C.POINTER(vp)]
        L.bcmi_free.argtypes = [vp]
        L.bcmi_memcpy.argtypes = [vp, vp, vp, sz, i, i]
        L.bcmi_fill.argtypes = [vp, vp, vp, sz, sz]
        L.bcmi_event_create.argtypes = [C.POINTER(vp)]
        L.bcmi_event_destroy.argtypes = [vp]
        L.bcmi_event_record.argtypes = [vp, vp]
        L.bcmi_event_synchronize.argtypes = [vp]
        L.bcmi_event_elapsed_ms.argtypes = [vp, vp, C.POINTER(C.c_float)]
        L.bcmi_sort.argtypes = [vp, i, i, vp, vp, sz, sz, i]
        L.bcmi_sort_noalloc.argtypes = [vp, i, i, vp, vp, sz, sz, i]
        L.bcmi_scan.argtypes = [vp, i, i, vp, vp, sz, vp, i, i]
        L.bcmi_reduce.argtypes = [vp, i, i, vp, sz, vp, vp, C.POINTER(C.c_double), i]
        L.bcmi_reduce_by_key.argtypes = [vp, i, i, i, vp, vp, sz, vp, vp, vp, i]
        L.bcmi_transform.argtypes = [vp, i, i, vp, vp, vp, sz]
        L.bcmi_saxpy.argtypes = [vp, C.c_float, vp, vp, vp, sz]
        L.bcmi_segmented_sort.argtypes = [vp, i, vp, vp, sz, vp, sz, i]
        L.bcmi_segmented_plan_create.argtypes = [vp, i, sz, vp, sz, i, C.POINTER(vp)]
        L.bcmi_segmented_sort_plan.argtypes = [vp, vp, vp, vp]
        L.bcmi_segmented_plan_destroy.argtypes = [vp]
        L.bcmi_segmented_scan.argtypes = [vp, i, i, vp, vp, vp, sz, vp, i, i]
        L.bcmi_graph_begin.argtypes = [vp]
        L.bcmi_graph_end.argtypes = [vp, C.POINTER(vp)]
        L.bcmi_graph_launch.argtypes = [vp, vp]
        L.bcmi_graph_destroy.argtypes = [vp]
        L.bcmi_tune_get.argtypes = [C.c_char_p, C.c_char_p, C.c_uint, C.POINTER(C.c_uint)]
        L.bcmi_tune_set.argtypes = [C.c_char_p, C.c_char_p, C.c_uint]
        L.bcmi_tuning_cache_load.argtypes = [C.c_char_p]
        L.bcmi_tuning_cache_save.argtypes = [C.c_char_p]
        L.bcmi_profile_enable.argtypes = [vp, i]
        L.bcmi_profile_query.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.bcmi_profile_reset.argtypes = [vp]
        pp = C.POINTER(Predicate)
        L.bcmi_copy_if.argtypes = [vp, i, vp, vp, sz, pp, i, vp]
        L.bcmi_count_if.argtypes = [vp, i, vp, sz, pp, vp]
        L.bcmi_unique.argtypes = [vp, i, vp, vp, sz, vp]
        L.bcmi_partition.argtypes = [vp, i, vp, vp, vp, sz, pp, vp]
        L.bcmi_extract_keys.argtypes = [vp, i, i, vp, sz, sz, sz, vp]
        L.bcmi_merge_sort.argtypes = [vp, C.POINTER(Comparator), vp, sz, vp, sz, sz]
        L.bcmi_scan_iter.argtypes = [vp, i, i, i, C.POINTER(Iter), vp, sz, vp, i, i]
        L.bcmi_transform_iter.argtypes = [vp, i, i, i, C.POINTER(Iter), C.POINTER(Iter), vp, sz]
        L.bcmi_convert.argtypes = [vp, i, i, vp, vp, sz]
        L.bcmi_debug_fault_probe.argtypes = [vp, C.c_uint]
        L.bcmi_host_malloc.argtypes = [sz, C.POINTER(vp)]
        L.bcmi_host_free.argtypes = [vp]
        L.bcmi_transform_reduce.argtypes = [vp, i, C.POINTER(Seq), C.POINTER(Seq), sz, C.POINTER(Map), i, vp, vp,
                                            C.POINTER(C.c_double), i]
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().bcmi_last_error()
        raise BcmiError(status, msg.decode() if msg else "")
    return status


def device_count():
    n = C.c_int()
    check(lib().bcmi_device_count(C.byref(n)))
    return n.value


def device_info(dev=0):
    info = DeviceInfo()
    check(lib().bcmi_device_get_info(dev, C.byref(info)))
    return {"name": info.name.decode(), "arch": info.arch.decode(),
            "compute_units": info.compute_units, "wavefront_size": info.wavefront_size,
            "clock_khz": info.clock_khz, "global_mem_bytes": info.global_mem_bytes,
            "lds_bytes_per_block": info.lds_bytes_per_block}


class Queue:
    """An in-order queue (the reference's command_queue) on one device."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().bcmi_queue_create(device, C.byref(self.h)))

    def finish(self):
        check(lib().bcmi_queue_finish(self.h))

    def reserve(self, n, elem_bytes):
        check(lib().bcmi_queue_reserve(self.h, n, elem_bytes))

    def close(self):
        if self.h:
            lib().bcmi_queue_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # profiling (HIP events around each kernel launch on this queue's stream)
    def profile(self, on=True):
        check(lib().bcmi_profile_enable(self.h, int(on)))

    def profile_query(self, kernel):
        n = C.c_int()
        ms = C.c_double()
        check(lib().bcmi_profile_query(self.h, kernel.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def profile_reset(self):
        check(lib().bcmi_profile_reset(self.h))


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        check(lib().bcmi_event_create(C.byref(self.h)))

    def record(self, q):
        check(lib().bcmi_event_record(self.h, q.h))

    def synchronize(self):
        check(lib().bcmi_event_synchronize(self.h))

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(lib().bcmi_event_elapsed_ms(self.h, end.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().bcmi_event_destroy(self.h)
        except Exception:
            pass


class DeviceArray:
    """Device allocation holding `size` elements of `dtype` (like vector<T>)."""

    def __init__(self, size, dtype, queue=None):
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self.nbytes = self.size * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(lib().bcmi_malloc(max(self.nbytes, 1), C.byref(self.ptr)))
        self.queue = queue

    @classmethod
    def from_host(cls, a, queue):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, queue)
        if a.size:
            check(lib().bcmi_memcpy(queue.h, d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D, 1))
        return d

    def to_host(self, queue=None, count=None):
        q = queue or self.queue
        count = self.size if count is None else count
        out = np.empty(count, self.dtype)
        if count:
            check(lib().bcmi_memcpy(q.h, out.ctypes.data_as(C.c_void_p), self.ptr,
                                    count * self.dtype.itemsize, D2H, 1))
        return out

    def at(self, offset):
        """Element pointer (buffer_iterator index) as a raw address."""
        return C.c_void_p(self.ptr.value + offset * self.dtype.itemsize)

    def free(self):
        if self.ptr:
            lib().bcmi_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    return x


# ---------------------------------------------------------------- algorithms
def sort(q, keys, values=None, api=SORT, descending=False, n=None, key_ptr=None, val_ptr=None):
    n = keys.size if n is None else n
    vb = values.dtype.itemsize if values is not None else 0
    check(lib().bcmi_sort(q.h, api, DTYPE[keys.dtype], key_ptr or keys.ptr,
                          (val_ptr or values.ptr) if values is not None else None, vb, n,
                          int(descending)))


def scan(q, inp, out, exclusive, init=None, op=PLUS, precision=COMPENSATED, n=None):
    n = inp.size if n is None else n
    ini = None
    if init is not None:
        ini_arr = np.array([init], dtype=inp.dtype)
        ini = ini_arr.ctypes.data_as(C.c_void_p)
    check(lib().bcmi_scan(q.h, DTYPE[inp.dtype], op, _p(inp), _p(out), n, ini, int(exclusive),
                          precision))


def reduce(q, inp, out, op=PLUS, init=None, precision=COMPENSATED, n=None):
    """Device-side reduce into out (DeviceArray of >=1 element). Returns the FP64
    value before final rounding (blocking) for float sums."""
    n = inp.size if n is None else n
    ini = None
    if init is not None:
        ini_arr = np.array([init], dtype=inp.dtype)
        ini = ini_arr.ctypes.data_as(C.c_void_p)
    f64 = DeviceArray(1, np.float64, q)
    check(lib().bcmi_reduce(q.h, DTYPE[inp.dtype], op, _p(inp), n, ini, _p(out),
                            C.cast(f64.ptr, C.POINTER(C.c_double)), precision))
    return f64


def reduce_by_key(q, keys, vals, keys_out, vals_out, op=PLUS, precision=COMPENSATED):
    cnt = DeviceArray(1, np.uint64, q)
    check(lib().bcmi_reduce_by_key(q.h, DTYPE[keys.dtype], DTYPE[vals.dtype], op, keys.ptr,
                                   vals.ptr, keys.size, keys_out.ptr, vals_out.ptr, cnt.ptr,
                                   precision))
    return int(cnt.to_host(q)[0])


def segmented_scan(q, keys, vals, out, exclusive, init=None, precision=COMPENSATED):
    ini = None
    if init is not None:
        ini_arr = np.array([init], dtype=vals.dtype)
        ini = ini_arr.ctypes.data_as(C.c_void_p)
    check(lib().bcmi_segmented_scan(q.h, DTYPE[keys.dtype], DTYPE[vals.dtype], keys.ptr, vals.ptr,
                                    out.ptr, keys.size, ini, int(exclusive), precision))


def transform(q, a, b, out, op):
    check(lib().bcmi_transform(q.h, DTYPE[a.dtype], op, _p(a), _p(b), _p(out), a.size))


def saxpy(q, a, x, y, out):
    check(lib().bcmi_saxpy(q.h, float(a), x.ptr, y.ptr, out.ptr, x.size))


# ------------------------------------------------------------- selection
def _count_buf(q, count):
    return count if count is not None else DeviceArray(1, np.uint64, q)


def copy_if(q, inp, out, pred, out_kind=SELECT_VALUES, n=None, count=None, in_ptr=None, out_ptr=None):
    """Selected elements (or their indices) of inp into out; returns the
    count (blocking read), or the device count buffer when count is given."""
    n = inp.size if n is None else n
    cnt = _count_buf(q, count)
    check(lib().bcmi_copy_if(q.h, DTYPE[inp.dtype], in_ptr or _p(inp), out_ptr or _p(out), n,
                             C.byref(pred), out_kind, cnt.ptr))
    return cnt if count is not None else int(cnt.to_host(q)[0])


def count_if(q, inp, pred, n=None, in_ptr=None):
    n = inp.size if n is None else n
    cnt = DeviceArray(1, np.uint64, q)
    check(lib().bcmi_count_if(q.h, DTYPE[inp.dtype], in_ptr or _p(inp), n, C.byref(pred), cnt.ptr))
    return int(cnt.to_host(q)[0])


def unique(q, inp, out, n=None, count=None, in_ptr=None, out_ptr=None):
    n = inp.size if n is None else n
    cnt = _count_buf(q, count)
    check(lib().bcmi_unique(q.h, DTYPE[inp.dtype], in_ptr or _p(inp), out_ptr or _p(out), n, cnt.ptr))
    return cnt if count is not None else int(cnt.to_host(q)[0])


def partition(q, inp, out_true, out_false, pred, n=None, count=None):
    """partition_copy (out_false given) or stable_partition into out_true."""
    n = inp.size if n is None else n
    cnt = _count_buf(q, count)
    check(lib().bcmi_partition(q.h, DTYPE[inp.dtype], _p(inp), _p(out_true), _p(out_false), n,
                               C.byref(pred), cnt.ptr))
    return cnt if count is not None else int(cnt.to_host(q)[0])


# ---------------------------------------------- transform_reduce / inner_product
def _as_seq(x, at=0):
    if isinstance(x, Seq):
        return x
    return Seq(x.ptr.value + at * x.dtype.itemsize if at else x.ptr.value, 0)


def transform_reduce(q, dtype, a, b, n, m, out, reduce_op=PLUS, init=None, precision=COMPENSATED, f64=None):
    """out[0] = init op y_0 op ... op y_{n-1}, y = m(a[i], b[i]) (b None: unary).
    a and b are DeviceArrays or counting(start).  Returns the FP64 value
    (a DeviceArray of one double, f64 when given: nothing is allocated then,
    as graph capture needs) for float sums."""
    dt = np.dtype(dtype)
    sa = _as_seq(a)
    sb = _as_seq(b) if b is not None else None
    ini = None
    if init is not None:
        ini_arr = np.array([init], dt)
        ini = ini_arr.ctypes.data_as(C.c_void_p)
    if f64 is None:
        f64 = DeviceArray(1, np.float64, q)
    check(lib().bcmi_transform_reduce(q.h, DTYPE[dt], C.byref(sa), C.byref(sb) if sb is not None else None, n,
                                      C.byref(m), reduce_op, ini, _p(out),
                                      C.cast(f64.ptr, C.POINTER(C.c_double)), precision))
    return f64


def sort_by_transform(q, records, key_dtype, key_offset=0, fn="id", descending=False, n=None):
    """experimental::sort_by_transform: keys = f(field at key_offset of each
    record) into a temporary, then sort_by_key(keys, records) (records of
    1, 2, 4, 8 or 16 bytes, in place)."""
    n = records.size if n is None else n
    kd = np.dtype(key_dtype)
    keys = DeviceArray(max(n, 1), kd, q)
    rb = records.dtype.itemsize
    check(lib().bcmi_extract_keys(q.h, DTYPE[kd], MAP_FN[fn], records.ptr, rb, key_offset, n, keys.ptr))
    check(lib().bcmi_sort(q.h, SORT_BY_KEY, DTYPE[kd], keys.ptr, records.ptr, rb, n, int(descending)))
    return keys


def merge_sort(q, cmp, records, record_bytes, n=None, values=None, value_bytes=None):
    """Comparison sort of n records of record_bytes bytes by a comparator
    (sort / stable_sort / sort_by_key with a user comparator), in place and
    stable; values (optional) move with their records."""
    n = records.nbytes // record_bytes if n is None else n
    if values is not None and value_bytes is None:
        value_bytes = values.nbytes // max(n, 1)
    check(lib().bcmi_merge_sort(q.h, C.byref(cmp), records.ptr, record_bytes,
                                values.ptr if values is not None else None, value_bytes or 0, n))


def segmented_sort(q, keys, values, offsets_host, descending=False):
    """Sort every segment [off[s], off[s+1]) as one stable_sort_by_key call
    would (blocking)."""
    off = np.ascontiguousarray(offsets_host, np.uint64)
    vb = values.dtype.itemsize if values is not None else 0
    check(lib().bcmi_segmented_sort(q.h, DTYPE[keys.dtype], keys.ptr,
                                    values.ptr if values is not None else None, vb,
                                    off.ctypes.data_as(C.c_void_p), off.size - 1, int(descending)))


class SegmentedPlan:
    """A segmented-sort plan: routing tables and temporaries on the device,
    built once; run() does no host work and allocates nothing (capturable)."""

    def __init__(self, q, key_dtype, value_dtype, offsets_host, descending=False):
        off = np.ascontiguousarray(offsets_host, np.uint64)
        vb = np.dtype(value_dtype).itemsize if value_dtype is not None else 0
        self.h = C.c_void_p()
        check(lib().bcmi_segmented_plan_create(q.h, DTYPE[np.dtype(key_dtype)], vb,
                                               off.ctypes.data_as(C.c_void_p), off.size - 1,
                                               int(descending), C.byref(self.h)))

    def run(self, q, keys, values=None):
        check(lib().bcmi_segmented_sort_plan(q.h, self.h, keys.ptr,
                                             values.ptr if values is not None else None))

    def __del__(self):
        try:
            lib().bcmi_segmented_plan_destroy(self.h)
        except Exception:
            pass


class Graph:
    """hipGraph capture/replay of work enqueued on a queue."""

    def __init__(self, q):
        self.q = q
        self.h = C.c_void_p()

    def __enter__(self):
        check(lib().bcmi_graph_begin(self.q.h))
        return self

    def __exit__(self, et, ev, tb):
        check(lib().bcmi_graph_end(self.q.h, C.byref(self.h)))
        return False

    def launch(self, q=None):
        check(lib().bcmi_graph_launch(self.h, (q or self.q).h))

    def __del__(self):
        try:
            lib().bcmi_graph_destroy(self.h)
        except Exception:
            pass
