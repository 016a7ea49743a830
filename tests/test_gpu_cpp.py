"""The header-only boost::compute templates driven from C++ on the GPU:
tests/cpp/test_shim.cpp runs the reference's own test cases (sort, radix
sort on a sub-range, sort_by_key, stable_sort_by_key, scans, partial_sum,
reduce, accumulate, reduce_by_key, transform, the saxpy lambda) through
the templates and libbcmi.so; it exits with the number of failed checks."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "build", "test_shim")


@pytest.mark.gpu
def test_cpp_shim_reference_cases():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/build/test_shim missing: run __graft_entry__.build() first")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


ITER_EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "build", "test_iterators")


@pytest.mark.gpu
def test_cpp_fancy_iterators_reference_cases():
    """tests/cpp/test_iterators.cpp: counting / constant / transform / zip
    iterators and int2_ / float4_ elements through the templates, on the
    reference's test vectors (test_scan.cpp, test_transform_reduce.cpp,
    test_reduce.cpp, test_zip_iterator.cpp, test_constant_iterator.cpp)."""
    if not os.path.exists(ITER_EXE):
        pytest.fail("tests/cpp/build/test_iterators missing: run __graft_entry__.build() first")
    r = subprocess.run([ITER_EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
