"""Writes reference_vectors.json: the hot-path golden vectors that Boost.Compute
0.5.0's own unit tests hold, transcribed as data (inputs and expected outputs).

Each case cites the reference test (path:line relative to the reference root)
it comes from.  "cmp": "eq" marks expectations the reference compares with
operator== (BOOST_CHECK_EQUAL_COLLECTIONS, test/check_macros.hpp:19-31), which
cannot tell -0.0 from +0.0; those cases are checked by value equality, and the
bit-level order is checked against the oracle separately.

Run:  python tests/golden/make_reference_vectors.py
"""
import json
import os

CASES = []


def case(cid, src, algo, dtype, **kw):
    d = {"id": cid, "source": src, "algo": algo, "dtype": dtype}
    d.update(kw)
    CASES.append(d)


# ------------------------------------------------------------ radix_sort
CHARS = [ord(c) for c in "ca07BF"] + [0, ord("$")]
radix_vectors = [
    ("char", "int8", CHARS, [0, ord("$"), ord("0"), ord("7"), ord("B"), ord("F"), ord("a"), ord("c")],
     [ord("c"), ord("a"), ord("F"), ord("B"), ord("7"), ord("0"), ord("$"), 0], 41, 237),
    ("uchar", "uint8", [0x12, 0x00, 0xFF, 0xB4, 0x80, 0x32, 0x64, 0xA2],
     [0x00, 0x12, 0x32, 0x64, 0x80, 0xA2, 0xB4, 0xFF],
     [0xFF, 0xB4, 0xA2, 0x80, 0x64, 0x32, 0x12, 0x00], 62, 270),
    ("short", "int16", [-4, 152, -94, 963, 31002, -456, 0, -2113],
     [-2113, -456, -94, -4, 0, 152, 963, 31002],
     [31002, 963, 152, 0, -4, -94, -456, -2113], 79, 303),
    ("ushort", "uint16", [4, 152, 94, 963, 63202, 34560, 0, 2113],
     [0, 4, 94, 152, 963, 2113, 34560, 63202],
     [63202, 34560, 2113, 963, 152, 94, 4, 0], 96, 336),
    ("int", "int32", [-4, 152, -5000, 963, 75321, -456, 0, 1112],
     [-5000, -456, -4, 0, 152, 963, 1112, 75321],
     [75321, 1112, 963, 152, 0, -4, -456, -5000], 113, 369),
    ("uint", "uint32", [500, 1988, 123456, 562, 0, 4000000, 9852, 102030],
     [0, 500, 562, 1988, 9852, 102030, 123456, 4000000],
     [4000000, 123456, 102030, 9852, 1988, 562, 500, 0], 128, 402),
    ("long", "int64", [500, 1988, 123456, 562, 0, 4000000, 9852, 102030],
     [0, 500, 562, 1988, 9852, 102030, 123456, 4000000], None, 145, None),
    ("ulong", "uint64", [500, 1988, 123456, 562, 0, 4000000, 9852, 102030],
     [0, 500, 562, 1988, 9852, 102030, 123456, 4000000],
     [4000000, 123456, 102030, 9852, 1988, 562, 500, 0], 162, 468),
]
for name, dt, inp, asc, desc, line_a, line_d in radix_vectors:
    case("radix_sort_%s" % name, "test/test_radix_sort.cpp:%d" % line_a, "radix_sort", dt,
         input=inp, expected=asc, descending=False)
    if desc is not None:
        case("radix_sort_%s_desc" % name, "test/test_radix_sort.cpp:%d" % line_d, "radix_sort", dt,
             input=inp, expected=desc, descending=True)
case("radix_sort_long_desc", "test/test_radix_sort.cpp:435", "radix_sort", "int64",
     input=[-500, 1988, 123456, 562, 0, 4000000, 9852, 102030],
     expected=[4000000, 123456, 102030, 9852, 1988, 562, 0, -500], descending=True)
FLOATS = [-6023.0, 152.5, -63.0, 1234567.0, 11.2, -5000.1, 0.0, 14.0, -8.25, -0.0]
case("radix_sort_float", "test/test_radix_sort.cpp:179", "radix_sort", "float32", input=FLOATS,
     expected=[-6023.0, -5000.1, -63.0, -8.25, -0.0, 0.0, 11.2, 14.0, 152.5, 1234567.0],
     descending=False, cmp="eq")
case("radix_sort_float_desc", "test/test_radix_sort.cpp:481", "radix_sort", "float32", input=FLOATS,
     expected=[1234567.0, 152.5, 14.0, 11.2, 0.0, -0.0, -8.25, -63.0, -5000.1, -6023.0],
     descending=True, cmp="eq")
case("radix_sort_double", "test/test_radix_sort.cpp:205", "radix_sort", "float64", input=FLOATS,
     expected=[-6023.0, -5000.1, -63.0, -8.25, -0.0, 0.0, 11.2, 14.0, 152.5, 1234567.0],
     descending=False, cmp="eq")
case("radix_sort_double_desc", "test/test_radix_sort.cpp:511", "radix_sort", "float64", input=FLOATS,
     expected=[1234567.0, 152.5, 14.0, 11.2, 0.0, -0.0, -8.25, -63.0, -5000.1, -6023.0],
     descending=True, cmp="eq")
# partial range vec.begin()+2 .. end()-2 (test/test_radix_sort.cpp:530-540)
case("radix_sort_partial_range", "test/test_radix_sort.cpp:530", "radix_sort_range", "int32",
     input=[9, 8, 7, 6, 5, 4, 3, 2, 1, 0], first=2, last=8,
     expected=[9, 8, 2, 3, 4, 5, 6, 7, 1, 0], descending=False)

# ------------------------------------------------------ radix_sort_by_key
KI = [10, 9, 2, 7, 6, -1, 4, 2, 2, 10]
VI = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10]
case("radix_sort_by_key_int_int", "test/test_radix_sort_by_key.cpp:40", "radix_sort_by_key", "int32",
     vdtype="int32", keys=KI, values=VI, expected_keys=[-1, 2, 2, 2, 4, 6, 7, 9, 10, 10],
     expected_values=[6, 3, 8, 9, 7, 5, 4, 2, 1, 10], descending=False)
case("radix_sort_by_key_int_int_desc", "test/test_radix_sort_by_key.cpp:69", "radix_sort_by_key",
     "int32", vdtype="int32", keys=KI, values=VI, expected_keys=[10, 10, 9, 7, 6, 4, 2, 2, 2, -1],
     expected_values=[1, 10, 2, 4, 5, 7, 3, 8, 9, 6], descending=True)
KU = [10, 9, 2, 7, 6, 1, 4, 2, 2, 10]
case("radix_sort_by_key_uint_uint", "test/test_radix_sort_by_key.cpp:108", "radix_sort_by_key",
     "uint32", vdtype="uint32", keys=KU, values=VI, expected_keys=[1, 2, 2, 2, 4, 6, 7, 9, 10, 10],
     expected_values=[6, 3, 8, 9, 7, 5, 4, 2, 1, 10], descending=False)
case("radix_sort_by_key_uint_uint_desc", "test/test_radix_sort_by_key.cpp:137", "radix_sort_by_key",
     "uint32", vdtype="uint32", keys=KU, values=VI, expected_keys=[10, 10, 9, 7, 6, 4, 2, 2, 2, 1],
     expected_values=[1, 10, 2, 4, 5, 7, 3, 8, 9, 6], descending=True)
KF = [10.0, 5.5, 10.0, 7.0, 5.5]
VF = [1, 200, -10, 2, 4]
case("radix_sort_by_key_float_int", "test/test_radix_sort_by_key.cpp:177", "radix_sort_by_key",
     "float32", vdtype="int32", keys=KF, values=VF, expected_keys=[5.5, 5.5, 7.0, 10.0, 10.0],
     expected_values=[200, 4, 2, 1, -10], descending=False)
case("radix_sort_by_key_float_int_desc", "test/test_radix_sort_by_key.cpp:206", "radix_sort_by_key",
     "float32", vdtype="int32", keys=KF, values=VF, expected_keys=[10.0, 10.0, 7.0, 5.5, 5.5],
     expected_values=[1, -10, 2, 200, 4], descending=True)
case("radix_sort_by_key_int_char", "test/test_radix_sort_by_key.cpp:246", "radix_sort_by_key", "int32",
     vdtype="int8", keys=[6, 1, 1, 3, 4, 7, 5, 1], values=[ord(c) for c in "gcbdehfa"],
     expected_keys=[1, 1, 1, 3, 4, 5, 6, 7], expected_values=[ord(c) for c in "cbadefgh"],
     descending=False)
# int2_ values (8 bytes): (1, i) packed as int64 with the pair's memory layout
int2 = [(1, i) for i in range(1, 9)]
packed = [((y & 0xFFFFFFFF) << 32) | (x & 0xFFFFFFFF) for x, y in int2]
order = [1, 2, 7, 3, 4, 6, 0, 5]
case("radix_sort_by_key_int_int2", "test/test_radix_sort_by_key.cpp:273", "radix_sort_by_key", "int32",
     vdtype="uint64", keys=[6, 1, 1, 3, 4, 7, 5, 1], values=packed,
     expected_keys=[1, 1, 1, 3, 4, 5, 6, 7], expected_values=[packed[i] for i in order],
     descending=False)

# -------------------------------------------------------------- public sort
case("sort_int_2", "test/test_sort.cpp:52", "sort", "int32", input=[4, 2], expected=[2, 4])
case("sort_float_3", "test/test_sort.cpp:71", "sort", "float32", input=[2.3, 0.1, 1.2],
     expected=[0.1, 1.2, 2.3])
for name, dt, inp, asc, desc, line_a, line_d in radix_vectors:
    line = {"char": 79, "uchar": 93, "short": 107, "ushort": 121, "int": 135, "uint": 147,
            "long": 161, "ulong": 175}[name]
    case("sort_%s" % name, "test/test_sort.cpp:%d" % line, "sort", dt, input=inp, expected=asc)
case("sort_float", "test/test_sort.cpp:189", "sort", "float32", input=FLOATS,
     expected=[-6023.0, -5000.1, -63.0, -8.25, -0.0, 0.0, 11.2, 14.0, 152.5, 1234567.0], cmp="eq")
case("sort_double", "test/test_sort.cpp:205", "sort", "float64", input=FLOATS,
     expected=[-6023.0, -5000.1, -63.0, -8.25, -0.0, 0.0, 11.2, 14.0, 152.5, 1234567.0], cmp="eq")
case("reverse_sort_int", "test/test_sort.cpp:227", "sort", "int32",
     input=[-4, 152, -5000, 963, 75321, -456, 0, 1112],
     expected=[75321, 1112, 963, 152, 0, -4, -456, -5000], descending=True)
case("sort_host_vector", "test/test_sort.cpp:285", "sort", "int32", input=[5, 2, 3, 6, 7, 4, 0, 1],
     expected=[0, 1, 2, 3, 4, 5, 6, 7])

# ------------------------------------------------------------ sort_by_key
case("sort_by_key_int_2", "test/test_sort_by_key.cpp:56", "sort_by_key", "int32", vdtype="int32",
     keys=[4, 2], values=[42, 24], expected_keys=[2, 4], expected_values=[24, 42])
case("sort_by_key_char_by_int", "test/test_sort_by_key.cpp:70", "sort_by_key", "int32", vdtype="int8",
     keys=[6, 2, 1, 3, 4, 7, 5, 0], values=[ord(c) for c in "gcbdehfa"],
     expected_keys=list(range(8)), expected_values=[ord(c) for c in "abcdefgh"])
n = 1024
case("sort_by_key_int_and_float", "test/test_sort_by_key.cpp:82", "sort_by_key", "int32",
     vdtype="float32", keys=[n - i for i in range(n)], values=[(n - i) / 2.0 for i in range(n)],
     expected_keys=[i + 1 for i in range(n)], expected_values=[(i + 1) / 2.0 for i in range(n)])
# stable_sort_by_key tie order (test/test_stable_sort_by_key.cpp:165-204)
size = 128
k = [-i for i in range(size)]
v = [-i for i in range(size)]
k[size // 2] = -256
k[size - 2] = -256
k[size - 1] = -256
v[size // 2] = 3
v[size - 2] = 1
v[size - 1] = 2
case("stable_sort_by_key_mid", "test/test_stable_sort_by_key.cpp:165", "stable_sort_by_key_prefix",
     "int32", vdtype="int32", keys=k, values=v, expected_keys_prefix=[-256, -256, -256],
     expected_values_prefix=[3, 1, 2])

# ------------------------------------------------------------------- scans
d12 = list(range(12))
case("inclusive_scan_int", "test/test_scan.cpp:41", "inclusive_scan", "int32", input=d12,
     expected=[0, 1, 3, 6, 10, 15, 21, 28, 36, 45, 55, 66])
case("exclusive_scan_int", "test/test_scan.cpp:93", "exclusive_scan", "int32", input=d12, init=0,
     expected=[0, 0, 1, 3, 6, 10, 15, 21, 28, 36, 45, 55])
case("inclusive_scan_int_multiplies", "test/test_scan.cpp:329", "inclusive_scan", "int32",
     input=[1, 2, 1, 2, 3], op="multiplies", expected=[1, 2, 2, 4, 12])
case("exclusive_scan_int_multiplies", "test/test_scan.cpp:359", "exclusive_scan", "int32",
     input=[1, 2, 1, 2, 3], init=10, op="multiplies", expected=[10, 10, 20, 20, 40])


def wrap32(x):
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= (1 << 31) else x


acc, inc = 1, []
for _ in range(1000):
    acc = wrap32(acc * 2)
    inc.append(acc)
case("inclusive_scan_int_multiplies_long", "test/test_scan.cpp:392", "inclusive_scan", "int32",
     input=[2] * 1000, op="multiplies", expected=inc)
acc, exc = 10, []
for _ in range(1000):
    exc.append(acc)
    acc = wrap32(acc * 2)
case("exclusive_scan_int_multiplies_long", "test/test_scan.cpp:413", "exclusive_scan", "int32",
     input=[2] * 1000, init=10, op="multiplies", expected=exc)
case("partial_sum_int", "test/test_partial_sum.cpp:28", "inclusive_scan", "int32",
     input=[1, 2, 5, 3, 9, 1, 4, 2], expected=[1, 3, 8, 11, 20, 21, 25, 27])

# ----------------------------------------------------------------- reduce
case("reduce_int_plus", "test/test_reduce.cpp:29", "reduce", "int32", input=[1, 5, 9, 13, 17],
     op="plus", expected=45)
case("reduce_int_multiplies", "test/test_reduce.cpp:29", "reduce", "int32", input=[1, 5, 9, 13, 17],
     op="multiplies", expected=9945)
case("reduce_doctest", "test/test_reduce.cpp:50", "reduce", "int32", input=[1, 2, 3, 4], op="plus",
     expected=10)
case("reduce_twos_plus", "test/test_reduce.cpp:62", "reduce", "uint32", input=[2] * 8, op="plus",
     expected=16)
case("reduce_twos_multiplies", "test/test_reduce.cpp:62", "reduce", "uint32", input=[2] * 8,
     op="multiplies", expected=256)
case("reduce_on_device_lo", "test/test_reduce.cpp:77", "reduce", "int32", input=[1, 2, 3, 4],
     op="plus", expected=10)
case("reduce_on_device_hi", "test/test_reduce.cpp:77", "reduce", "int32", input=[5, 6, 7, 8],
     op="plus", expected=26)
case("reduce_int_min", "test/test_reduce.cpp:88", "reduce", "int32", input=[11, 5, 92, 13, 42],
     op="min", expected=5)
case("reduce_int_max", "test/test_reduce.cpp:88", "reduce", "int32", input=[11, 5, 92, 13, 42],
     op="max", expected=92)

# ------------------------------------------------------------- accumulate
for init, exp in ((0, 20), (-10, 10), (5, 25)):
    case("accumulate_sum_int_init%d" % init, "test/test_accumulate.cpp:25", "accumulate", "int32",
         input=[2, 4, 6, 8], init=init, op="plus", expected=exp)
for init, exp in ((1, 384), (-1, -384), (2, 768)):
    case("accumulate_product_int_init%d" % init, "test/test_accumulate.cpp:42", "accumulate", "int32",
         input=[2, 4, 6, 8], init=init, op="multiplies", expected=exp)
case("accumulate_std_equality_int", "test/test_accumulate.cpp:268", "accumulate", "int32",
     input=[1, 2, 3, 4], init=0, op="plus", expected=10)
case("accumulate_std_equality_int_10000", "test/test_accumulate.cpp:272", "accumulate", "int32",
     input=[2] * 10000, init=0, op="plus", expected=20000)

# ----------------------------------------------------------- reduce_by_key
case("reduce_by_key_int", "test/test_reduce_by_key.cpp:26", "reduce_by_key", "int32", vdtype="int32",
     keys=[0, 2, -3, -3, -3, -3, -3, 4], values=[1, 1, 1, 1, 1, 2, 5, 1], op="plus",
     expected_keys=[0, 2, -3, 4], expected_values=[1, 1, 10, 1])
keys = [0] * 1024
for i in (137, 677, 1001):
    keys[i] = 1
run = 0
scanned = []
for x in keys:
    run += x
    scanned.append(run)
case("reduce_by_key_int_long_uniform", "test/test_reduce_by_key.cpp:48", "reduce_by_key", "int32",
     vdtype="int32", keys=[0] * 1024, values=[1] * 1024, op="plus", expected_keys=[0],
     expected_values=[1024])
case("reduce_by_key_int_long_runs", "test/test_reduce_by_key.cpp:62", "reduce_by_key", "int32",
     vdtype="int32", keys=scanned, values=[1] * 1024, op="plus", expected_keys=[0, 1, 2, 3],
     expected_values=[137, 540, 324, 23])
case("reduce_by_key_one", "test/test_reduce_by_key.cpp:88", "reduce_by_key", "int32", vdtype="int32",
     keys=[22], values=[-9], op="plus", expected_keys=[22], expected_values=[-9])
KM = [0, 2, 2, 3, 3, 3, 3, 3, 4]
case("reduce_by_key_int_min", "test/test_reduce_by_key.cpp:104", "reduce_by_key", "int32",
     vdtype="int32", keys=KM, values=[1, 2, 1, -3, 1, 4, 2, 5, 77], op="min",
     expected_keys=[0, 2, 3, 4], expected_values=[1, 1, -3, 77])
case("reduce_by_key_int_max", "test/test_reduce_by_key.cpp:104", "reduce_by_key", "int32",
     vdtype="int32", keys=KM, values=[1, 2, 1, -3, 1, 4, 2, 5, 77], op="max",
     expected_keys=[0, 2, 3, 4], expected_values=[1, 2, 5, 77])
case("reduce_by_key_float_max", "test/test_reduce_by_key.cpp:131", "reduce_by_key", "int32",
     vdtype="float32", keys=KM, values=[1.0, 2.0, -1.5, -3.0, 1.0, -0.24, 2, 5, 77.1], op="max",
     expected_keys=[0, 2, 3, 4], expected_values=[1.0, 2.0, 5.0, 77.1], rtol=1e-6)

# -------------------------------------------------------------- transform
case("transform_add_int", "test/test_transform.cpp:75", "transform", "int32", a=[1, 2, 3, 4],
     b=[10, 20, 30, 40], op="plus", expected=[11, 22, 33, 44])
case("transform_mul_int", "test/test_transform.cpp:75", "transform", "int32", a=[1, 2, 3, 4],
     b=[10, 20, 30, 40], op="multiplies", expected=[10, 40, 90, 160])

# -------------------------------------------------------------- selection
# Predicates: {"fn": id|mul|mod|and|signbit, "a": operand, "cmp": op, "c": rhs,
# "negate": not1}.  The vector-type cases (float4 dot product,
# test/test_copy_if.cpp:70; int4/int2 counts, test/test_count.cpp:74,159) and
# the constant_iterator case (test/test_count.cpp:44) are outside the
# scalar-element path and are not transcribed.
def pred(cmp, c, fn="id", a=0, negate=False):
    return {"fn": fn, "a": a, "cmp": cmp, "c": c, "negate": negate}


CI = [1, 6, 3, 5, 8, 2, 4]
case("copy_if_int_lt5", "test/test_copy_if.cpp:37", "copy_if", "int32", input=CI,
     pred=pred("<", 5), expected=[1, 3, 2, 4])
case("copy_if_int_x2_ge10", "test/test_copy_if.cpp:44", "copy_if", "int32", input=CI,
     pred=pred(">=", 10, "mul", 2), expected=[6, 5, 8])
ODD = [1, 2, 3, 4, 5, 1, 2, 3, 4, 5]
case("copy_if_odd", "test/test_copy_if.cpp:59", "copy_if", "int32", input=ODD,
     pred=pred("==", 1, "mod", 2), expected=[1, 3, 5, 1, 3, 5])
case("copy_if_even", "test/test_copy_if.cpp:66", "copy_if", "int32", input=ODD,
     pred=pred("==", 0, "mod", 2), expected=[2, 4, 2, 4])
case("copy_index_if_int", "test/test_copy_if.cpp:115", "copy_index_if", "int32", input=CI,
     pred=pred("<", 5), expected=[0, 2, 5, 6])

case("unique_int", "test/test_unique.cpp:35", "unique", "int32", input=[1, 6, 6, 4, 2, 2, 4],
     expected=[1, 6, 4, 2, 4, 2, 4], expected_count=5)
case("unique_all_same_float", "test/test_unique.cpp:47", "unique", "float32", input=[3.14] * 1024,
     expected_count=1, expected_prefix=[3.14])
IOTA = list(range(1024))
case("unique_iota_uint", "test/test_unique.cpp:66", "unique", "uint32", input=IOTA, expected_count=1024,
     expected=IOTA)
EVENS = [x - (x & 1) for x in IOTA]
case("unique_even_uints", "test/test_unique.cpp:86", "unique", "uint32", input=EVENS, expected_count=512,
     expected_prefix=list(range(0, 1024, 2)))
case("unique_copy_int", "test/test_unique_copy.cpp:33", "unique_copy", "int32", input=[1, 6, 6, 4, 2, 2, 4],
     expected=[1, 6, 4, 2, 4])

# partition() is stable_partition() (partition.hpp); the reference test checks
# the count and the signs of the two groups
case("partition_float_signbit", "test/test_partition.cpp:51", "partition", "float32",
     input=[1.0, 2.0, -1.0, -2.0, 3.0, 4.0, -3.0, -4.0], pred=pred("!=", 0, "signbit"),
     expected=[-1.0, -2.0, -3.0, -4.0, 1.0, 2.0, 3.0, 4.0], expected_count=4)
case("partition_small_one", "test/test_partition.cpp:75", "partition", "float32", input=[1.0],
     pred=pred("!=", 0, "signbit"), expected=[1.0], expected_count=0)
case("partition_small_two", "test/test_partition.cpp:79", "partition", "float32", input=[1.0, -1.0],
     pred=pred("!=", 0, "signbit"), expected=[-1.0, 1.0], expected_count=1)
case("stable_partition_int", "test/test_stable_partition.cpp:34", "partition", "int32",
     input=[1, 1, -2, 0, 5, -1, 2, 4, 0, -1], pred=pred(">", 0),
     expected=[1, 1, 5, 2, 4, -2, 0, -1, 0, -1], expected_count=5)

# remove(v) = remove_if(_1 == v) = copy_if(not1(_1 == v)) back into the range
# (remove.hpp, remove_if.hpp); each step works on the prefix the last left
case("remove_int", "test/test_remove.cpp:32", "remove", "int32", input=[1, 2, 1, 3, 2, 4, 3, 4, 5],
     steps=[{"len": 9, "value": 2, "count": 7, "expected": [1, 1, 3, 4, 3, 4, 5]},
            {"len": 7, "value": 4, "count": 5, "expected": [1, 1, 3, 3, 5]},
            {"len": 5, "value": 1, "count": 3, "expected": [3, 3, 5]},
            {"len": 3, "value": 5, "count": 2, "expected": [3, 3]},
            {"len": 2, "value": 3, "count": 0, "expected": []}])

CNT = [1, 2, 1, 2, 3]
for line, lo, hi, v, want in [(36, 0, 5, 1, 2), (37, 0, 5, 2, 2), (38, 0, 5, 3, 1), (39, 1, 5, 1, 1),
                              (40, 1, 4, 3, 0), (41, 1, 4, 2, 2)]:
    case("count_int_%d" % line, "test/test_count.cpp:%d" % line, "count_if", "int32", input=CNT[lo:hi],
         pred=pred("==", v), expected=want)
case("count_if_greater_than_two", "test/test_count.cpp:69", "count_if", "float32",
     input=[1.0, 2.5, -1.0, 3.0, 5.0], pred=pred(">", 2.0), expected=3)
NL = [ord(ch) for ch in "abcdefg\nhijklmn\nopqrs\ntuv\nwxyz\n"]
case("count_newlines", "test/test_count.cpp:124", "count_if", "int8", input=NL, pred=pred("==", 10), expected=5)
UC = [0x00, 0x10, 0x2F, 0x10, 0x01, 0x00, 0x01, 0x00]
for line, v, want in [(138, 0x00, 3), (142, 0x10, 2), (146, 0x2F, 1), (150, 0x01, 2), (154, 0xFF, 0)]:
    case("count_uchar_%d" % line, "test/test_count.cpp:%d" % line, "count_if", "uint8", input=UC,
         pred=pred("==", v), expected=want)
IOTA2K = list(range(2048))
case("count_if_odd", "test/test_count.cpp:202", "count_if", "int32", input=IOTA2K,
     pred=pred("!=", 0, "and", 1), expected=1024)
case("count_if_with_reduce", "test/test_count.cpp:217", "count_if", "int32", input=IOTA2K,
     pred=pred(">", 1024), expected=1023)

# ----------------------------------------------- transform_reduce / inner_product
# {"op": binary combine, "fn": g, "c": unary offset}; a "count" input is
# make_counting_iterator(start).  The float4 length case
# (test/test_transform_reduce.cpp:42) is on a vector type and is not transcribed.
case("transform_reduce_sum_abs_int", "test/test_transform_reduce.cpp:39", "transform_reduce", "int32",
     a=[1, -2, -3, -4, 5], map={"fn": "abs", "c": 0}, reduce="plus", expected=15)
case("transform_reduce_variance", "test/test_transform_reduce.cpp:88", "transform_reduce", "float32",
     a=[1, 2, 3, 4, 5, 6, 7, 8, 9, 10], map={"fn": "square", "c": 5.5}, reduce="plus", expected=82.5, rtol=1e-4)
case("inner_product_int", "test/test_inner_product.cpp:36", "inner_product", "int32", a=[1, 2, 3, 4],
     b=[10, 20, 30, 40], init=0, expected=300)
case("inner_product_counting_iterator", "test/test_inner_product.cpp:41", "inner_product", "int32",
     a={"count": 0, "n": 100}, b={"count": 0}, init=0, expected=328350)

# ------------------------------------------------ experimental::sort_by_transform
# records as lists of components; "get" = component index (None: the scalar
# itself), "fn" = id | abs.  The length<float4_> case
# (test/test_sort_by_transform.cpp:41) has implementation-defined rounding and
# is not transcribed.
case("sort_int_by_abs", "test/test_sort_by_transform.cpp:25", "sort_by_transform", "int32",
     records=[1, -2, 4, -3, 0, 5, -8, -9], get=None, fn="abs", descending=False,
     expected=[0, 1, -2, -3, 4, 5, -8, -9])
case("sort_float4_by_component", "test/test_sort_by_transform.cpp:75", "sort_by_transform", "float32",
     records=[[1, 2, 3, 0], [9, 8, 7, 0], [4, 5, 6, 0], [0, 0, 0, 0]], get=1, fn="id", descending=False,
     expected=[[0, 0, 0, 0], [1, 2, 3, 0], [4, 5, 6, 0], [9, 8, 7, 0]])

# ------------------------------------------------------ comparator sorts
# sort / stable_sort / sort_by_key / stable_sort_by_key with a user
# comparator (merge_sort_on_gpu).  records: lists of components (or scalars);
# cmp = {"form": field | abs | abs_sum2 | length, "lhs"/"rhs": component read
# from the left / right operand, "second": the second addend (abs_sum2),
# "count": components (length), "greater": CMP is >}.  The checks are the
# test's own: "expected" (whole output), "expected_x" (positions -> first
# component), "expected_prefix" (+ "expected_values_prefix"), and "sorted"
# (is_sorted by the comparator, plus values sorted where the test checks it).
def _gen(n, f):
    return [f(i) for i in range(n)]


case("merge_sort_gpu_char_less", "test/test_merge_sort_gpu.cpp:36-53", "compare_sort", "int8",
     records=[99, 97, 48, 55, 66, 70, 0, 36], cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": False},
     expected=[0, 36, 48, 55, 66, 70, 97, 99])
case("merge_sort_gpu_char_greater", "test/test_merge_sort_gpu.cpp:55-62", "compare_sort", "int8",
     records=[0, 36, 48, 55, 66, 70, 97, 99], cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": True},
     expected=[99, 97, 70, 66, 55, 48, 36, 0])
case("merge_sort_gpu_int_abs", "test/test_merge_sort_gpu.cpp:224-258", "compare_sort", "int32",
     records=_gen(1024, lambda i: 1024 - i if i % 2 else i - 1024), cmp={"form": "abs", "lhs": 0, "rhs": 0,
                                                                         "greater": False}, sorted=True)
case("merge_sort_gpu_int2_abs_sum", "test/test_merge_sort_gpu.cpp:261-296", "compare_sort", "int32",
     records=_gen(1024, lambda i: [i, i] if i % 2 else [i - 1024, i - 1024]),
     cmp={"form": "abs_sum2", "lhs": 0, "rhs": 0, "second": 1, "greater": False}, sorted=True)
case("merge_sort_gpu_long8_s0_s3", "test/test_merge_sort_gpu.cpp:299-330", "compare_sort", "int64",
     records=_gen(256, lambda i: [i] * 8 if i % 2 else [i * i] * 8),
     cmp={"form": "field", "lhs": 0, "rhs": 3, "greater": False}, sorted=True)
case("merge_sort_gpu_stable_int2", "test/test_merge_sort_gpu.cpp:333-370", "compare_sort", "int32",
     records=[[8, 3], [5, 1], [2, 1], [6, 1], [8, 1], [7, 1], [4, 1], [8, 2]],
     cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": False}, stable=True,
     expected=[[2, 1], [4, 1], [5, 1], [6, 1], [7, 1], [8, 3], [8, 1], [8, 2]])
case("stable_sort_int2_first", "test/test_stable_sort.cpp:43-73", "compare_sort", "int32",
     records=[[2, 1], [2, 2], [1, 2], [1, 1]], cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": False},
     stable=True, expected=[[1, 2], [1, 1], [2, 1], [2, 2]])
case("stable_sort_int2_second", "test/test_stable_sort.cpp:75-92", "compare_sort", "int32",
     records=[[1, 2], [1, 1], [2, 1], [2, 2]], cmp={"form": "field", "lhs": 1, "rhs": 1, "greater": False},
     stable=True, expected=[[1, 1], [2, 1], [1, 2], [2, 2]])
case("sort_float2_by_length", "test/test_sort.cpp:238-285", "compare_sort", "float32",
     records=[[1.0, 0.2], [1.3, 1.0], [6.7, 0.0], [5.2, 3.4], [1.4, 1.4]],
     cmp={"form": "length", "lhs": 0, "rhs": 0, "count": 2, "greater": False},
     expected=[[1.0, 0.2], [1.3, 1.0], [1.4, 1.4], [5.2, 3.4], [6.7, 0.0]])
case("sort_particles_by_x", "test/test_sort.cpp:293-326", "compare_sort", "float32",
     records=[[0.1, 0.0], [-0.4, 0.0], [10.0, 0.0], [0.001, 0.0]],
     cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": False},
     expected=[[-0.4, 0.0], [0.001, 0.0], [0.1, 0.0], [10.0, 0.0]])
_int2 = [[0, 0]] * 100
_int2 = [list(r) for r in _int2]
_int2[0], _int2[25], _int2[75], _int2[97], _int2[51] = [100, 0], [20, 0], [9, 0], [-10, 0], [-10, -1]
case("sort_int2_by_x", "test/test_sort.cpp:328-365", "compare_sort", "int32", records=_int2,
     cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": False},
     expected_x={"0": -10, "1": -10, "97": 9, "98": 20, "99": 100}, sorted=True)
case("sort_by_key_int_custom_less", "test/test_sort_by_key.cpp:113-142", "compare_sort", "int32",
     records=_gen(1024, lambda i: 1024 - i), values=_gen(1024, lambda i: (1024 - i) / 2.0), values_dtype="float32",
     cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": False}, sorted=True, values_sorted=True)
_sk = _gen(128, lambda i: -i)
_sv = _gen(128, lambda i: -i)
_sk[64], _sk[126], _sk[127] = -256, -256, -256
_sv[64], _sv[126], _sv[127] = 3, 1, 2
case("stable_sort_by_key_int_custom_less", "test/test_stable_sort_by_key.cpp:165-208", "compare_sort", "int32",
     records=_sk, values=_sv, values_dtype="int32", cmp={"form": "field", "lhs": 0, "rhs": 0, "greater": False},
     stable=True, sorted=True, expected_prefix=[-256, -256, -256], expected_values_prefix=[3, 1, 2])

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.json")
    with open(out, "w") as f:
        json.dump({"reference": "Boost.Compute 0.5.0", "cases": CASES}, f, indent=1)
    print("wrote %d cases to %s" % (len(CASES), out))
