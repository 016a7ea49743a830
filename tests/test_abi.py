"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point include/bcmi.h declares, and its host-side logic (errors, device
selection, tuning cache) behaves.  No kernel is launched here."""
import ctypes as C
import os
import re
import subprocess

import pytest

import compute_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "bcmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bcmi_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    names = declared_entry_points()
    assert len(names) >= 35
    L = ca.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.check_output(["nm", "-D", "--defined-only", ca.LIB_PATH], text=True)
    exported = set(re.findall(r" T (bcmi_\w+)", out))
    assert set(names) <= exported
    assert set(names) == set(ca.EXPORTS)


def test_no_gpu_path_fails_loudly():
    n = ca.device_count()
    with pytest.raises(ca.BcmiError) as e:
        ca.Queue(n + 3)  # a device that does not exist
    assert e.value.code == -1  # CL_DEVICE_NOT_FOUND
    assert "not found" in str(e.value)


def test_null_arguments_are_rejected():
    L = ca.lib()
    assert L.bcmi_device_count(None) == -30
    assert L.bcmi_queue_finish(None) == -36
    assert L.bcmi_malloc(16, None) == -30
    assert b"null" in L.bcmi_last_error()


def test_tuning_cache_round_trip(tmp_path):
    L = ca.lib()
    ca.check(L.bcmi_tune_set(b"__bcmi_radix_sort_uint", b"ipt", 12))
    v = C.c_uint()
    ca.check(L.bcmi_tune_get(b"__bcmi_radix_sort_uint", b"ipt", 16, C.byref(v)))
    assert v.value == 12
    ca.check(L.bcmi_tune_get(b"__bcmi_radix_sort_uint", b"tpb", 256, C.byref(v)))
    assert v.value == 256
    path = str(tmp_path / "tune.json").encode()
    ca.check(L.bcmi_tuning_cache_save(path))
    ca.check(L.bcmi_tune_set(b"__bcmi_radix_sort_uint", b"ipt", 8))
    ca.check(L.bcmi_tuning_cache_load(path))
    ca.check(L.bcmi_tune_get(b"__bcmi_radix_sort_uint", b"ipt", 16, C.byref(v)))
    assert v.value == 12
    assert L.bcmi_tuning_cache_load(str(tmp_path / "missing.json").encode()) == -30


def test_version():
    a, b = C.c_int(), C.c_int()
    ca.check(ca.lib().bcmi_get_version(C.byref(a), C.byref(b)))
    assert (a.value, b.value) == (0, 5)


CPP_BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "build")


def test_cpp_headers_link_in_two_translation_units():
    """The header-only boost::compute templates, included by two objects of
    one program, link without duplicate symbols (the reference's
    test/extra/test_multiple_objects1.cpp / test_multiple_objects2.cpp)."""
    exe = os.path.join(CPP_BUILD, "odr")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp not built (run __graft_entry__.build())")
    assert subprocess.run([exe], timeout=60).returncode == 0


def test_cpp_headers_compile_standalone():
    """Every public header compiles on its own (include-what-you-use)."""
    inc = os.path.join(os.path.dirname(CPP_BUILD), "..", "..", "include")
    inc = os.path.normpath(inc)
    headers = []
    for d, _, files in os.walk(os.path.join(inc, "boost")):
        for f in files:
            if f.endswith(".hpp"):
                headers.append(os.path.relpath(os.path.join(d, f), inc))
    assert len(headers) > 20
    src = "".join('#include <%s>\n' % h for h in sorted(headers))
    for h in sorted(headers):
        r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I", inc, "-x", "c++", "-"],
                           input="#include <%s>\nint main(){return 0;}\n" % h, text=True,
                           capture_output=True, timeout=120)
        assert r.returncode == 0, (h, r.stderr[:2000])
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I", inc, "-x", "c++", "-"],
                       input=src + "int main(){return 0;}\n", text=True, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr[:2000]


@pytest.mark.parametrize("expr,ok", [
    ("bc::_1 < 5", True), ("bc::_1 * 2 >= 10", True), ("bc::_1 % 2 == 1", True), ("(bc::_1 & 1) != 0", True),
    ("bc::_1 & 1", True), ("bc::not1(bc::_1 > 0)", True),
    # C would compare an int element with an unsigned or double literal in
    # that wider type; the kernels compare in int, so these must not compile
    ("bc::_1 < 5u", False), ("bc::_1 > 2.5", False), ("bc::_1 * 2.0 > 1", False),
    ("bc::signbit_<float>()", False), ("bc::_1 + 1", False),
])
def test_selection_predicates_follow_c_conversions_at_compile_time(expr, ok):
    """A predicate literal whose type would make C compare in another type
    than the kernel does is rejected when the program is compiled."""
    inc = os.path.normpath(os.path.join(os.path.dirname(CPP_BUILD), "..", "..", "include"))
    src = ("#include <boost/compute/algorithm/count_if.hpp>\nnamespace bc = boost::compute;\n"
           "size_t f(bc::vector<int> &v, bc::command_queue &q) { return bc::count_if(v.begin(), v.end(), %s, q); }\n"
           "int main(){return 0;}\n" % expr)
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I", inc, "-x", "c++", "-"], input=src, text=True,
                       capture_output=True, timeout=120)
    assert (r.returncode == 0) == ok, r.stderr[:1500]
