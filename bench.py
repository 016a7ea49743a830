#!/usr/bin/env python3
"""Benchmark of the data-parallel primitive core on one MI355X.

Headline (the `value`): radix sort of 2^27 uint32 keys (uniform over 32 bits,
mt19937 seed 0xB0057001), ascending, through the C-ABI, inputs resident in
HBM.  One step = one sort of fresh (unsorted) keys: an untimed device-to-device
copy restores the input, then the sort is timed with HIP events on the
queue's stream.  The same JSON line carries:
  * roofline: the dominant kernel (one onesweep pass) -- algorithmic bytes
    per launch (8 B per key: read + write) over its average launch duration,
    measured with HIP events around each launch on the queue's stream;
    traffic from the committed rocprofv3 PMC summary when present;
  * cpu_baseline: the oracle's restatement of the reference's CPU-device sort
    (merge_sort_on_cpu) timed on a bounded sample on this host;
  * secondary: scan / reduce / saxpy / sort_by_key / reduce_by_key /
    segmented scan / segmented sort at their configured sizes, each with GB/s
    and fraction of the 8 TB/s HBM peak, plus the float error of the
    compensated reduce against FP64.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--only a,b,..]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBS = 8000.0  # MI355X HBM3E
SECTIONS = ["sort", "sort_by_key", "scan", "reduce", "saxpy", "rbk", "segscan", "select", "mapreduce", "msort",
            "segsort"]


def timed_steps(ca, q, fn, steps, warmup, restore=None):
    for _ in range(warmup):
        if restore:
            restore()
        fn()
    q.finish()
    evs = [(ca.Event(), ca.Event()) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in evs:
        if restore:
            restore()
        s.record(q)
        fn()
        e.record(q)
    q.finish()
    wall = time.perf_counter() - t0
    ms = [s.elapsed_ms(e) for s, e in evs]
    return ms, wall


def kernel_avg_ms(q, fn, kernel, reps=3, restore=None):
    q.profile_reset()
    q.profile(True)
    for _ in range(reps):
        if restore:
            restore()
        fn()
    q.profile(False)
    n, tot = q.profile_query(kernel)
    q.profile_reset()
    return (tot / n if n else float("nan")), n // reps if reps else n


def kernel_breakdown(q, fn, kernels, reps=3, restore=None):
    """Per-call milliseconds of each named kernel (queue event profiler)."""
    q.profile_reset()
    q.profile(True)
    for _ in range(reps):
        if restore:
            restore()
        fn()
    q.profile(False)
    out = {}
    for k in kernels:
        n, tot = q.profile_query(k)
        if n:
            out[k] = round(tot / reps, 4)
    q.profile_reset()
    return out


def d2d(ca, q, dst, src):
    ca.check(ca.lib().bcmi_memcpy(q.h, dst.ptr, src.ptr, src.nbytes, ca.D2D, 0))


def gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The reference's CPU-device column (what Boost.Compute runs on a CPU
    OpenCL device such as PoCL), restated on OpenMP threads in
    oracle/cpu_column.c, timed per primitive on bounded samples of the same
    workloads: one untimed warm-up, then the minimum of 3 trials
    (perf/perf.hpp conventions).  Neither Boost nor an OpenCL CPU device is
    on this machine (the probe is kept with the profiles), so the reference
    itself cannot run here."""
    import oracle as O
    import compute_amd.workloads as W
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)

    def best(fn, trials=3):
        fn()
        ts = []
        for _ in range(trials):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    per = {}
    m = 1 << 24
    keys = W.sort_u32(m)
    t_sort = best(lambda: O.cc_sort(keys, threads=threads))
    per["sort_u32"] = {"Gkeys_per_s": round(m / t_sort / 1e9, 5), "sample": "2^24 keys", "s": round(t_sort, 3)}
    mk = 1 << 22
    k, v = W.sort_by_key_f32(mk)
    t = best(lambda: O.cc_sort(k, v, threads=threads))
    per["sort_by_key_f32_u32"] = {"Gpairs_per_s": round(mk / t / 1e9, 5), "sample": "2^22 pairs",
                                  "s": round(t, 3)}
    ms = 1 << 26
    x = W.scan_i32(ms)
    t = best(lambda: O.cc_scan_i32(x, True, 0, threads=threads))
    per["exclusive_scan_i32"] = {"GBps": round(8 * ms / t / 1e9, 2), "sample": "2^26 elements", "s": round(t, 4)}
    f = W.lognormal_f32(ms)
    t = best(lambda: O.cc_reduce_f32(f, threads=threads))
    per["reduce_f32"] = {"GBps": round(4 * ms / t / 1e9, 2), "sample": "2^26 elements", "s": round(t, 4)}
    t = best(lambda: O.cc_accumulate_f32(f, 0.0), trials=1)
    per["accumulate_f32_serial_fold"] = {"GBps": round(4 * ms / t / 1e9, 2), "sample": "2^26 elements",
                                         "threads": 1, "s": round(t, 4)}
    a, b = W.saxpy_f32(ms)
    t = best(lambda: O.cc_saxpy(2.5, a, b, threads=threads))
    per["saxpy_f32"] = {"GBps": round(12 * ms / t / 1e9, 2), "sample": "2^26 elements", "s": round(t, 4)}
    mr = 1 << 25
    rk, rv = W.runs_i32_f32(mr)
    t = best(lambda: O.cc_reduce_by_key(rk, rv), trials=1)
    per["reduce_by_key_i32_f32_serial"] = {"GBps": round(8 * mr / t / 1e9, 2), "sample": "2^25 elements",
                                           "threads": 1, "s": round(t, 4)}
    off = W.segments(65536)
    nseg = 65536
    sub = off
    sk, sv = W.segment_data(int(sub[-1]))
    t = best(lambda: O.cc_segmented_sort(sk, sv, sub, threads=threads), trials=1)
    per["segmented_sort_per_call_loop"] = {
        "Gpairs_per_s": round(int(sub[-1]) / t / 1e9, 5), "s": round(t, 3),
        "sample": "all %d segments (%d pairs), one merge_sort_by_key_on_cpu call per segment, "
                  "each on all threads" % (nseg, int(sub[-1]))}
    del keys, k, v, x, f, a, b, rk, rv, sk, sv
    # divergence A5: the reference's float accumulate is one serial FP32
    # fold; its error on the full 2^28 precision workload, against the
    # double-double sum
    full = W.lognormal_f32(1 << 28)
    fold = float(O.cc_accumulate_f32(full, 0.0))
    ref = W.sum_dd(full)
    ulp = float(np.spacing(np.float32(abs(ref))))
    per["accumulate_f32_serial_fold"]["err_on_2p28_vs_fp64"] = {
        "ulp": round(abs(fold - ref) / ulp, 3), "abs": float("%.4g" % abs(fold - ref)),
        "rel": float("%.4g" % (abs(fold - ref) / abs(ref)))}
    return {"value": round(m / t_sort / 1e9, 5), "unit": "Gkeys/s", "cores": threads, "kind": "port",
            "sample": "first 2^24 of the same mt19937(0xB0057001) keys, sorted by oracle/cpu_column.c's "
                      "merge_sort_on_cpu on %d OpenMP threads (min of 3 after a warm-up); %.3f s" % (threads, t_sort),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "per_primitive": per}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--only", default="all")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-call", action="store_true",
                    help="skip the segmented sort's one-call-per-segment baseline (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: the library has no multi-GPU path "
              "(each GPU would sort or scan its own shard); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    only = SECTIONS if args.only == "all" else args.only.split(",")

    import compute_amd as ca
    import compute_amd.workloads as W

    q = ca.Queue(0)
    info = ca.device_info(0)
    K, Wm = args.steps, args.warmup
    sec = {}
    out = None

    # ---------------------------------------------------- headline: sort u32
    if "sort" in only:
        n = 1 << 27
        keys = W.sort_u32(n)
        src = ca.DeviceArray.from_host(keys, q)
        work = ca.DeviceArray(n, np.uint32, q)
        restore = lambda: d2d(ca, q, work, src)
        fn = lambda: ca.sort(q, work)
        ms, wall = timed_steps(ca, q, fn, K, Wm, restore)
        got = work.to_host(q)
        assert np.all(got[1:] >= got[:-1]), "sort output not sorted"
        ms_step = float(np.mean(ms))
        os_ms, launches = kernel_avg_ms(q, fn, "radix_onesweep", restore=restore)
        h_ms, _ = kernel_avg_ms(q, fn, "radix_histogram", restore=restore)
        per_launch = 8 * n  # one pass: read 4 B + write 4 B per key
        ach = gbs(per_launch, os_ms)
        tr = load_traffic("radix_onesweep_u32_2p27_traffic.json")
        out = {
            "metric": "sort Gkeys/s & scan/reduce GB/s (%HBM peak) at 2^27; batched vs per-call",
            "value": round(n / (ms_step * 1e-3) / 1e9, 3),
            "unit": "Gkeys/s",
            "n_gpus": 1, "steps": K, "warmup": Wm,
            "ms_per_step": round(ms_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32",
            "data": "synthetic: 2^27 uint32 keys = raw mt19937(0xB0057001) output",
            "config": {"workload": "radix sort of 2^27 uint32 keys, uniform over 32 bits, ascending "
                                   "(boost::compute::sort / detail::radix_sort)",
                       "n": n, "device": info["name"], "arch": info["arch"],
                       "timing": "hipEvent around each sort call on the queue stream; input restored "
                                 "by an untimed D2D copy before each step",
                       "wall_ms_per_step_incl_restore": round(wall * 1e3 / K, 4),
                       "alg_bytes_per_key": 36,
                       "sort_GBps_alg": round(gbs(36 * n, ms_step), 1),
                       "sort_frac_of_hbm": round(gbs(36 * n, ms_step) / PEAK_GBS, 4),
                       "kernels_ms": {"radix_histogram": round(h_ms, 4), "radix_onesweep_avg": round(os_ms, 4),
                                      "onesweep_launches_per_sort": launches}},
            "roofline": {"bound": "hbm", "kernel": "radix_onesweep (one 8-bit digit pass)",
                         "achieved": round(ach, 1), "peak": PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / PEAK_GBS, 4),
                         "alg_bytes_per_launch": per_launch,
                         "traffic": tr.get("bytes_per_launch") if tr else None},
        }
        del src, work, keys, got

    # ------------------------------------------------------- sort_by_key f32
    if "sort_by_key" in only:
        n = 1 << 27
        k, v = W.sort_by_key_f32(n)
        sk, sv = ca.DeviceArray.from_host(k, q), ca.DeviceArray.from_host(v, q)
        wk, wv = ca.DeviceArray(n, np.float32, q), ca.DeviceArray(n, np.uint32, q)

        def restore():
            d2d(ca, q, wk, sk)
            d2d(ca, q, wv, sv)

        fn = lambda: ca.sort(q, wk, wv, api=ca.SORT_BY_KEY)
        ms, _ = timed_steps(ca, q, fn, max(2, K // 2), 1, restore)
        t = float(np.mean(ms))
        os_ms, _ = kernel_avg_ms(q, fn, "radix_onesweep", restore=restore)
        sec["sort_by_key_f32_u32_2p27"] = {
            "Gpairs_per_s": round(n / (t * 1e-3) / 1e9, 3), "ms": round(t, 4),
            "GBps_alg_68B": round(gbs(68 * n, t), 1), "frac": round(gbs(68 * n, t) / PEAK_GBS, 4),
            "onesweep_pass_GBps": round(gbs(16 * n, os_ms), 1)}
        del sk, sv, wk, wv, k, v

    # ------------------------------------------------------ exclusive scan
    if "scan" in only:
        n = 1 << 28
        x = W.scan_i32(n)
        dx = ca.DeviceArray.from_host(x, q)
        do = ca.DeviceArray(n, np.int32, q)
        fn = lambda: ca.scan(q, dx, do, True, init=0)
        ms, _ = timed_steps(ca, q, fn, K, Wm)
        t = float(np.mean(ms))
        kms, _ = kernel_avg_ms(q, fn, "scan")
        sec["exclusive_scan_i32_2p28"] = {"GBps": round(gbs(8 * n, t), 1), "ms": round(t, 4),
                                          "frac": round(gbs(8 * n, t) / PEAK_GBS, 4),
                                          "kernel_ms": round(kms, 4)}
        del dx, do, x

    # --------------------------------------------- reduce / inclusive scan
    if "reduce" in only:
        n = 1 << 28
        x = W.lognormal_f32(n)
        dx = ca.DeviceArray.from_host(x, q)
        r = ca.DeviceArray(1, np.float32, q)
        fn = lambda: ca.reduce(q, dx, r)
        ms, _ = timed_steps(ca, q, fn, K, Wm)
        t = float(np.mean(ms))
        got = float(r.to_host(q)[0])
        ca.reduce(q, dx, r, precision=ca.NAIVE)
        naive = float(r.to_host(q)[0])
        # errors against the double-double value of the sum (workloads.py)
        ref = W.sum_dd(x)
        ulp = float(np.spacing(np.float32(abs(ref))))
        ca.reduce(q, dx, r, init=np.float32(1.5))
        acc = float(r.to_host(q)[0])
        ref_acc = W.sum_dd(x, 1.5)
        do = ca.DeviceArray(n, np.float32, q)
        fs = lambda: ca.scan(q, dx, do, False)
        ms2, _ = timed_steps(ca, q, fs, max(2, K // 2), 1)
        t2 = float(np.mean(ms2))
        scan_err = W.err_scan(x, do.to_host(q), False)
        ca.scan(q, dx, do, False, precision=ca.NAIVE)
        scan_err_naive = W.err_scan(x, do.to_host(q), False)
        rnd = lambda e: {k: float("%.4g" % v) for k, v in e.items()}
        sec["reduce_f32_lognormal_2p28"] = {
            "GBps": round(gbs(4 * n, t), 1), "ms": round(t, 4), "frac": round(gbs(4 * n, t) / PEAK_GBS, 4),
            "fp64_ref": "double-double sum",
            "compensated_err_ulp_vs_fp64": round(abs(got - ref) / ulp, 3),
            "compensated_err_abs": float("%.4g" % abs(got - ref)),
            "compensated_err_rel": float("%.4g" % (abs(got - ref) / abs(ref))),
            "naive_fp32_err_ulp_vs_fp64": round(abs(naive - ref) / ulp, 3),
            "naive_fp32_err_abs": float("%.4g" % abs(naive - ref)),
            "naive_fp32_err_rel": float("%.4g" % (abs(naive - ref) / abs(ref))),
            "accumulate_init_1p5_err_ulp_vs_fp64": round(abs(acc - ref_acc) /
                                                         float(np.spacing(np.float32(abs(ref_acc)))), 3),
            "inclusive_scan_GBps": round(gbs(8 * n, t2), 1), "inclusive_scan_ms": round(t2, 4),
            "inclusive_scan_frac": round(gbs(8 * n, t2) / PEAK_GBS, 4),
            "inclusive_scan_err_vs_fp64_compensated": rnd(scan_err),
            "inclusive_scan_err_vs_fp64_naive": rnd(scan_err_naive)}
        del dx, do, x

    # ---------------------------------------------------------------- saxpy
    if "saxpy" in only:
        n = 1 << 28
        x, y = W.saxpy_f32(n)
        dx, dy = ca.DeviceArray.from_host(x, q), ca.DeviceArray.from_host(y, q)
        do = ca.DeviceArray(n, np.float32, q)
        fn = lambda: ca.saxpy(q, 2.5, dx, dy, do)
        ms, _ = timed_steps(ca, q, fn, K, Wm)
        t = float(np.mean(ms))
        # FP64 reference: 2.5*x is exact in double, so one double rounding
        # separates it from the exact a*x + y
        ref = 2.5 * x.astype(np.float64) + y.astype(np.float64)
        err = np.abs(do.to_host(q).astype(np.float64) - ref)
        ulps = err / np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
        sec["saxpy_f32_2p28"] = {"GBps": round(gbs(12 * n, t), 1), "ms": round(t, 4),
                                 "frac": round(gbs(12 * n, t) / PEAK_GBS, 4),
                                 "err_vs_fp64": {"max_abs": float("%.4g" % err.max()),
                                                 "max_ulp": float("%.4g" % ulps.max())}}
        del ref, err, ulps
        del dx, dy, do, x, y

    # -------------------------------------------- reduce_by_key / seg scan
    if "rbk" in only or "segscan" in only:
        n = 1 << 27
        k, v = W.runs_i32_f32(n)
        dk, dv = ca.DeviceArray.from_host(k, q), ca.DeviceArray.from_host(v, q)
        ko, vo = ca.DeviceArray(n, np.int32, q), ca.DeviceArray(n, np.float32, q)
        cnt = ca.DeviceArray(1, np.uint64, q)
        fn = lambda: ca.check(ca.lib().bcmi_reduce_by_key(q.h, ca.I32, ca.F32, ca.PLUS, dk.ptr, dv.ptr, n,
                                                          ko.ptr, vo.ptr, cnt.ptr, ca.COMPENSATED))
        ms, _ = timed_steps(ca, q, fn, K, Wm)
        t = float(np.mean(ms))
        runs = int(cnt.to_host(q)[0])
        alg = 8 * n + 8 * runs
        rnd = lambda e: {k: float("%.4g" % v) for k, v in e.items()}
        assert np.array_equal(ko.to_host(q, count=runs), np.arange(runs, dtype=np.int32)), "reduce_by_key keys"
        sec["reduce_by_key_i32_f32_2p27"] = {"GBps": round(gbs(alg, t), 1), "ms": round(t, 4),
                                             "frac": round(gbs(alg, t) / PEAK_GBS, 4), "runs": runs,
                                             "run_sum_err_vs_fp64": rnd(W.err_runsum(k, v, vo.to_host(q, count=runs)))}
        fs = lambda: ca.segmented_scan(q, dk, dv, vo, True, init=0.0)
        ms, _ = timed_steps(ca, q, fs, K, Wm)
        t = float(np.mean(ms))
        sec["segmented_exclusive_scan_2p27"] = {"GBps": round(gbs(12 * n, t), 1), "ms": round(t, 4),
                                                "frac": round(gbs(12 * n, t) / PEAK_GBS, 4),
                                                "err_vs_fp64": rnd(W.err_segscan(k, v, vo.to_host(q), True))}
        del dk, dv, ko, vo, k, v

    # -------------------------------------------------- stream compaction
    if "select" in only:
        import ctypes
        n = 1 << 28
        x = W.scan_i32(n)  # the exclusive-scan input: int32 uniform in [-100, 100]
        dx = ca.DeviceArray.from_host(x, q)
        do, df = ca.DeviceArray(n, np.int32, q), ca.DeviceArray(n, np.int32, q)
        cnt = ca.DeviceArray(1, np.uint64, q)
        pr = ca.predicate(np.int32, ">=", 0)
        kept = int(np.count_nonzero(x >= 0))
        fn = lambda: ca.copy_if(q, dx, do, pr, count=cnt)
        ms, _ = timed_steps(ca, q, fn, K, Wm)
        t = float(np.mean(ms))
        assert int(cnt.to_host(q)[0]) == kept, "copy_if count"
        kern = kernel_breakdown(q, fn, ["select"])
        alg = 4 * n + 4 * kept
        sec["copy_if_i32_2p28"] = {"GBps": round(gbs(alg, t), 1), "ms": round(t, 4),
                                   "frac": round(gbs(alg, t) / PEAK_GBS, 4), "selected": kept,
                                   "alg_bytes": "4 B read per element + 4 B per selected element",
                                   "kernels_ms": kern}
        fc = lambda: ca.check(ca.lib().bcmi_count_if(q.h, ca.I32, dx.ptr, n, ctypes.byref(pr), cnt.ptr))
        ms, _ = timed_steps(ca, q, fc, K, Wm)
        t = float(np.mean(ms))
        assert int(cnt.to_host(q)[0]) == kept, "count_if count"
        sec["count_if_i32_2p28"] = {"GBps": round(gbs(4 * n, t), 1), "ms": round(t, 4),
                                    "frac": round(gbs(4 * n, t) / PEAK_GBS, 4)}
        fp = lambda: ca.partition(q, dx, do, df, pr, count=cnt)
        ms, _ = timed_steps(ca, q, fp, K, Wm)
        t = float(np.mean(ms))
        sec["partition_copy_i32_2p28"] = {"GBps": round(gbs(8 * n, t), 1), "ms": round(t, 4),
                                          "frac": round(gbs(8 * n, t) / PEAK_GBS, 4)}
        fs = lambda: ca.partition(q, dx, do, None, pr, count=cnt)
        ms, _ = timed_steps(ca, q, fs, K, Wm)
        t = float(np.mean(ms))
        got = do.to_host(q)
        assert np.array_equal(got[:kept], x[x >= 0]) and np.array_equal(got[kept:], x[x < 0]), "stable_partition"
        sec["stable_partition_i32_2p28"] = {"GBps_vs_8B": round(gbs(8 * n, t), 1), "ms": round(t, 4),
                                            "frac_vs_8B": round(gbs(8 * n, t) / PEAK_GBS, 4),
                                            "passes": 2, "note": "two passes (16 B/element moved); "
                                                                 "frac against the 8 B/element minimum"}
        del dx, do, df, x, got
        nk = 1 << 27
        k, _ = W.runs_i32_f32(nk)
        dk, dko = ca.DeviceArray.from_host(k, q), ca.DeviceArray(nk, np.int32, q)
        fu = lambda: ca.unique(q, dk, dko, count=cnt)
        ms, _ = timed_steps(ca, q, fu, K, Wm)
        t = float(np.mean(ms))
        runs = int(cnt.to_host(q)[0])
        alg = 4 * nk + 4 * runs
        sec["unique_i32_runs_2p27"] = {"GBps": round(gbs(alg, t), 1), "ms": round(t, 4),
                                       "frac": round(gbs(alg, t) / PEAK_GBS, 4), "runs": runs}
        del dk, dko, k

    # ------------------------------------------- transform_reduce / inner_product
    if "mapreduce" in only:
        n = 1 << 28
        x, y = W.saxpy_f32(n)  # the saxpy inputs: float32 ~ N(0, 1)
        dx, dy = ca.DeviceArray.from_host(x, q), ca.DeviceArray.from_host(y, q)
        res1 = ca.DeviceArray(1, np.float32, q)
        mp = ca.mapping(np.float32, "id", "multiplies")
        fi = lambda: ca.transform_reduce(q, np.float32, dx, dy, n, mp, res1)
        ms, _ = timed_steps(ca, q, fi, K, Wm)
        t = float(np.mean(ms))
        got = float(res1.to_host(q)[0])
        ref = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
        ulp = float(np.spacing(np.float32(abs(ref))))
        kern = kernel_breakdown(q, fi, ["mapreduce", "mapreduce_final"])
        sec["inner_product_f32_2p28"] = {"GBps": round(gbs(8 * n, t), 1), "ms": round(t, 4),
                                         "frac": round(gbs(8 * n, t) / PEAK_GBS, 4),
                                         "err_ulp_vs_fp64": round(abs(got - ref) / ulp, 3), "kernels_ms": kern}
        mean = float(np.mean(x, dtype=np.float64))
        mv = ca.mapping(np.float32, "square", c=np.float32(mean))
        fv = lambda: ca.transform_reduce(q, np.float32, dx, None, n, mv, res1)
        ms, _ = timed_steps(ca, q, fv, K, Wm)
        t = float(np.mean(ms))
        got = float(res1.to_host(q)[0])
        c32 = float(np.float32(mean))
        ref = float(np.sum(np.square(x.astype(np.float64) - c32)))
        sec["sum_sq_dev_f32_2p28"] = {"GBps": round(gbs(4 * n, t), 1), "ms": round(t, 4),
                                      "frac": round(gbs(4 * n, t) / PEAK_GBS, 4),
                                      "err_ulp_vs_fp64": round(abs(got - ref) / float(np.spacing(np.float32(ref))), 3)}
        del dx, dy, x, y, res1

    # ------------------------------------------- comparison (merge) sort
    if "msort" in only:
        # 2^24 int2_ records sorted by get<0>(_1) < get<0>(_2), stable: x
        # uniform in [0, 2^20) (ties), y = the input position
        n = 1 << 24
        rng = np.random.default_rng(0xB0057001)
        rec = np.stack([rng.integers(0, 1 << 20, n), np.arange(n)], axis=1).astype(np.int32)
        src, work = ca.DeviceArray.from_host(rec, q), ca.DeviceArray(2 * n, np.int32, q)
        cmp = ca.comparator(np.int32, "field", 0)
        fn = lambda: ca.merge_sort(q, cmp, work, 8, n)
        restore = lambda: d2d(ca, q, work, src)
        # default route: the key codes sorted by radix, then one gather
        ms, _ = timed_steps(ca, q, fn, K, Wm, restore)
        t = float(np.mean(ms))
        kern_key = kernel_breakdown(q, fn, ["merge_sort_codes", "radix_histogram", "radix_onesweep",
                                            "merge_sort_gather"], restore=restore)
        # the merge kernels (the route of comparators reading different fields)
        ca.check(ca.lib().bcmi_tune_set(b"__bcmi_merge_sort", b"route", 0))
        ms, _ = timed_steps(ca, q, fn, K, Wm, restore)
        t_merge = float(np.mean(ms))
        kern = kernel_breakdown(q, fn, ["merge_sort_block", "merge_sort_split", "merge_sort_merge"], restore=restore)
        ca.check(ca.lib().bcmi_tune_set(b"__bcmi_merge_sort", b"route", 1))
        # the sort_by_key form: uint32 values ride along (codes + (code, index) radix sort + one gather)
        vsrc, vwork = ca.DeviceArray.from_host(np.arange(n, dtype=np.uint32), q), ca.DeviceArray(n, np.uint32, q)
        fkv = lambda: ca.merge_sort(q, cmp, work, 8, n, vwork, 4)
        restore_kv = lambda: (d2d(ca, q, work, src), d2d(ca, q, vwork, vsrc))
        ms, _ = timed_steps(ca, q, fkv, K, Wm, restore_kv)
        t_kv = float(np.mean(ms))
        kern_kv = kernel_breakdown(q, fkv, ["merge_sort_codes", "radix_histogram", "radix_onesweep",
                                            "merge_sort_gather"], restore=restore_kv)
        del vsrc, vwork
        passes = int(np.ceil(np.log2(n / 4096)))
        merge_ms = kern.get("merge_sort_merge", float("nan")) / passes
        sec["merge_sort_int2_by_x_2p24"] = {
            "Mrecords_per_s": round(n / (t * 1e-3) / 1e6, 1), "ms": round(t, 4), "kernels_ms": kern_key,
            "by_key_u32_values_ms": round(t_kv, 4), "by_key_kernels_ms": kern_kv,
            "merge_kernels_ms": round(t_merge, 4), "merge_kernels_Mrecords_per_s": round(n / (t_merge * 1e-3) / 1e6, 1),
            "merge_kernels_breakdown_ms": kern, "merge_passes": passes, "merge_pass_ms": round(merge_ms, 4),
            "merge_pass_GBps": round(gbs(16 * n, merge_ms), 1),
            "merge_pass_frac": round(gbs(16 * n, merge_ms) / PEAK_GBS, 4),
            "note": "default: key codes + stable radix sort_by_key of (code, index) + gather; merge kernels: "
                    "8-byte {code, index} items, a merge pass reads and writes 16 B per record, the last pass "
                    "gathers the records"}
        del src, work, rec

    # ------------------------------------------------------ segmented sort
    if "segsort" in only:
        off = W.segments(65536)
        n = int(off[-1])
        k, v = W.segment_data(n)
        sk, sv = ca.DeviceArray.from_host(k, q), ca.DeviceArray.from_host(v, q)
        wk, wv = ca.DeviceArray(n, np.uint64, q), ca.DeviceArray(n, np.uint32, q)

        def restore():
            d2d(ca, q, wk, sk)
            d2d(ca, q, wv, sv)

        plan = ca.SegmentedPlan(q, np.uint64, np.uint32, off)
        fn = lambda: plan.run(q, wk, wv)
        ms, _ = timed_steps(ca, q, fn, K, Wm, restore)
        t_eager = float(np.mean(ms))
        g = ca.Graph(q)
        with g:
            plan.run(q, wk, wv)
        ms, _ = timed_steps(ca, q, lambda: g.launch(), K, Wm, restore)
        t_graph = float(np.mean(ms))
        got = wk.to_host(q)
        head = np.zeros(n, bool)
        head[off[:-1][np.diff(off) > 0].astype(np.int64)] = True
        assert np.all((got[1:] >= got[:-1]) | head[1:]), "segmented sort output not sorted per segment"
        # per-call baseline: the build's own stable_sort_by_key once per segment
        lens = np.diff(off).astype(np.int64)
        L = ca.lib()

        def per_call():
            for s_, ln in zip(off[:-1].tolist(), lens.tolist()):
                if ln > 1:
                    L.bcmi_sort(q.h, ca.STABLE_SORT_BY_KEY, ca.U64, wk.at(s_), wv.at(s_), 4, ln, 0)

        # the same loop issued from C: no interpreter time between calls
        per_call_c = lambda: ca.check(W.sort_per_segment(L.bcmi_sort, q.h, ca.STABLE_SORT_BY_KEY, ca.U64,
                                                         wk.ptr, 8, wv.ptr, 4, off))

        parts = kernel_breakdown(q, fn, ["segsort_small", "segsort_insertion", "segsort_histogram",
                                         "segsort_digit_offsets", "segsort_route", "segsort_onesweep_msd",
                                         "segsort_groups", "segsort_onesweep", "segsort_copyback"],
                                 restore=restore)
        if args.no_per_call:
            t_call = t_call_py = float("nan")
        else:
            ms, wall = timed_steps(ca, q, per_call_c, 2, 1, restore)
            t_call = float(min(ms))
            got_c = wk.to_host(q)
            assert np.array_equal(got_c, got), "per-call loop and batched plan disagree"
            ms, wall = timed_steps(ca, q, per_call, 1, 1, restore)
            t_call_py = float(ms[0])
        sec["segmented_sort_u64_u32_65536seg"] = {
            "pairs": n, "segments": 65536,
            "batched_graph_ms": round(t_graph, 4), "batched_eager_ms": round(t_eager, 4),
            "per_call_ms": round(t_call, 2) if t_call == t_call else None,
            "per_call_note": "one stable_sort_by_key call per segment, looped in C (workloads.cpp)",
            "per_call_python_loop_ms": round(t_call_py, 2) if t_call_py == t_call_py else None,
            "speedup_graph_vs_per_call": round(t_call / t_graph, 1) if t_call == t_call else None,
            "kernels_ms_per_call": parts,
            "Gpairs_per_s": round(n / (t_graph * 1e-3) / 1e9, 3),
            "GBps_alg_24B": round(gbs(24 * n, t_graph), 1),
            "frac_24B": round(gbs(24 * n, t_graph) / PEAK_GBS, 4)}
        del plan, g, sk, sv, wk, wv, k, v

    if out is None:
        out = {"metric": "sort Gkeys/s & scan/reduce GB/s (%HBM peak) at 2^27; batched vs per-call",
               "value": None, "unit": "Gkeys/s", "n_gpus": 1, "steps": K, "warmup": Wm,
               "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "u32", "data": "synthetic", "config": {"workload": "partial run: " + args.only}}
    out["secondary"] = sec

    # ---------------------------------------------------------- CPU baseline
    if not args.no_cpu_baseline and "sort" in only:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
