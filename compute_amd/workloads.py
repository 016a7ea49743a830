"""Seeded synthetic inputs for the benchmark configurations (numpy arrays),
from the host-only helper libbcmi_workloads.so (std::mt19937 engines with
explicit transforms of their raw output)."""
import ctypes as C
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbcmi_workloads.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_PATH)
        vp, sz = C.c_void_p, C.c_size_t
        L.wl_sort_u32.argtypes = [C.c_uint32, vp, sz]
        L.wl_sort_by_key_f32.argtypes = [vp, vp, sz]
        L.wl_scan_i32.argtypes = [vp, sz]
        L.wl_saxpy_f32.argtypes = [vp, vp, sz]
        L.wl_lognormal_f32.argtypes = [vp, sz]
        L.wl_segments.argtypes = [vp, sz, sz]
        L.wl_segments.restype = sz
        L.wl_segment_data.argtypes = [vp, vp, sz]
        L.wl_runs_i32_f32.argtypes = [vp, vp, sz, sz]
        L.wl_scan_f32_dd.argtypes = [vp, vp, sz]
        L.wl_runs_i32_f32.restype = sz
        dp = C.POINTER(C.c_double)
        L.wl_sum_f32_dd.argtypes = [vp, sz, C.c_double]
        L.wl_sum_f32_dd.restype = C.c_double
        L.wl_err_scan_f32.argtypes = [vp, vp, sz, C.c_int, C.c_double, dp]
        L.wl_err_segscan_f32.argtypes = [vp, vp, vp, sz, C.c_int, dp]
        L.wl_err_runsum_f32.argtypes = [vp, vp, sz, vp, sz, dp]
        L.wl_err_runsum_f32.restype = sz
        L.wl_sort_per_segment.argtypes = [vp, vp, C.c_int, C.c_int, vp, sz, vp, sz, vp, sz]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def sort_u32(n, seed=0xB0057001):
    k = np.empty(n, np.uint32)
    lib().wl_sort_u32(seed, _p(k), n)
    return k


def sort_by_key_f32(n):
    k, v = np.empty(n, np.float32), np.empty(n, np.uint32)
    lib().wl_sort_by_key_f32(_p(k), _p(v), n)
    return k, v


def scan_i32(n):
    x = np.empty(n, np.int32)
    lib().wl_scan_i32(_p(x), n)
    return x


def saxpy_f32(n):
    x, y = np.empty(n, np.float32), np.empty(n, np.float32)
    lib().wl_saxpy_f32(_p(x), _p(y), n)
    return x, y


def lognormal_f32(n):
    x = np.empty(n, np.float32)
    lib().wl_lognormal_f32(_p(x), n)
    return x


def segments(nseg, max_len=1 << 20):
    off = np.empty(nseg + 1, np.uint64)
    lib().wl_segments(_p(off), nseg, max_len)
    return off


def segment_data(n):
    k, v = np.empty(n, np.uint64), np.empty(n, np.uint32)
    lib().wl_segment_data(_p(k), _p(v), n)
    return k, v


def runs_i32_f32(n, max_run=4096):
    k, v = np.empty(n, np.int32), np.empty(n, np.float32)
    lib().wl_runs_i32_f32(_p(k), _p(v), n, max_run)
    return k, v


# ------------------------------------------------- float error measurement
def _err(fn, *args):
    e = (C.c_double * 3)()
    r = fn(*args, e)
    return {"max_abs": e[0], "max_rel": e[1], "max_ulp": e[2]}, r


def sum_dd(x, init=0.0):
    """Double-double value of init + sum(x) (float32 x)."""
    return lib().wl_sum_f32_dd(_p(np.ascontiguousarray(x, np.float32)), x.size, float(init))


def scan_dd(x):
    """Double-double inclusive prefix sums of float32 x, as float64."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.size, np.float64)
    lib().wl_scan_f32_dd(_p(x), _p(out), x.size)
    return out


def err_scan(x, got, exclusive, init=0.0):
    """Max abs / rel / float-ulp error of a float32 prefix sum."""
    return _err(lib().wl_err_scan_f32, _p(x), _p(got), x.size, int(exclusive), float(init))[0]


def err_segscan(keys, x, got, exclusive):
    return _err(lib().wl_err_segscan_f32, _p(keys), _p(x), _p(got), x.size, int(exclusive))[0]


def err_runsum(keys, x, got):
    e, runs = _err(lib().wl_err_runsum_f32, _p(keys), _p(x), x.size, _p(got), got.size)
    assert runs == got.size, (runs, got.size)
    return e


def sort_per_segment(sort_fn, queue_handle, api, key_dtype, keys_ptr, key_bytes, vals_ptr, value_bytes, offsets):
    """One library sort call per segment, looped in C; returns the status."""
    off = np.ascontiguousarray(offsets, np.uint64)
    return lib().wl_sort_per_segment(C.cast(sort_fn, C.c_void_p), queue_handle, api, key_dtype, keys_ptr,
                                     key_bytes, vals_ptr, value_bytes, _p(off), off.size - 1)
