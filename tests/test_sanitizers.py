"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only; GPU sanitizers are not available for these kernels):

  * the C oracle and the CPU-column baseline (oracle/Makefile target
    "sanitize") run the whole oracle and baseline test files with the ASan
    runtime preloaded into the Python process;
  * the segmented sort's host plan builder (compute_amd/csrc/segplan.hpp) is
    compiled with both sanitizers into a driver that builds plans for
    random and edge-case segment lists and checks every table invariant.

A sanitizer report aborts the child process, which fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(p) or not os.path.exists(p):
        pytest.skip("gcc sanitizer runtime %s not found" % name)
    return p


def test_oracle_and_baseline_under_asan_ubsan():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    env = dict(os.environ)
    env["BCMI_ORACLE_BUILD"] = os.path.join(ROOT, "oracle", "build", "sanitize")
    env["LD_PRELOAD"] = _runtime("libasan.so")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    # the child proves which libraries it loaded
    probe = ("import oracle as O, numpy as np; O.lib(); O.cpucol(); "
             "m = open('/proc/self/maps').read(); "
             "assert 'sanitize/liboracle.so' in m and 'sanitize/libcpucol.so' in m and 'libasan' in m; print('ok')")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "tests/test_oracle.py", "tests/test_cpu_column.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "passed" in r.stdout


def test_segmented_plan_builder_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "segplan_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
                           "-I" + os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "segplan_check.cpp")])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "0 failures" in r.stdout
