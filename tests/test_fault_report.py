"""CPU side of the fault path (no GPU): the fault word -> status mapping of
bcmi_queue_finish (compute_amd/csrc/fault.hpp) and the opencl_error the
templates raise for it, built with the host compiler under AddressSanitizer
and UndefinedBehaviorSanitizer; and the Python binding's error type.  The
path from a kernel's give-up to the host is exercised on the GPU by
tests/test_gpu_fault.py."""
import os
import subprocess

import compute_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fault_mapping_and_exception(tmp_path):
    exe = str(tmp_path / "fault_check")
    subprocess.check_call(["g++", "-std=c++14", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-I" + os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "fault_check.cpp")])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def test_python_error_carries_the_status():
    try:
        ca.check(ca.OUT_OF_RESOURCES)
    except ca.BcmiError as e:
        assert e.code == -5
    else:
        raise AssertionError("check(-5) did not raise")
    assert ca.check(0) == 0
