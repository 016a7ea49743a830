"""The benchmark's input generators (compute_amd/csrc/workloads.cpp, built on
std::mt19937) and the oracle's independent ones produce identical bits."""
import numpy as np

import compute_amd.workloads as W
import oracle as O


def test_generators_agree_with_oracle():
    n = 1 << 17
    assert np.array_equal(W.sort_u32(n), O.gen_sort_u32(n))
    a, b = W.sort_by_key_f32(n), O.gen_sort_by_key_f32(n)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[1], b[1])
    assert np.array_equal(W.scan_i32(n), O.gen_scan_i32(n))
    a, b = W.saxpy_f32(n), O.gen_saxpy_f32(n)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(W.lognormal_f32(n).view(np.uint32), O.gen_lognormal_f32(n).view(np.uint32))
    assert np.array_equal(W.segments(65536), O.gen_segments(65536))
    a, b = W.segment_data(n), O.gen_segment_data(n)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    a, b = W.runs_i32_f32(n), O.gen_runs_i32_f32(n)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
