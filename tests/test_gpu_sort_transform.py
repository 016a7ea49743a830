"""experimental::sort_by_transform through the C-ABI (bcmi_extract_keys +
bcmi_sort(SORT_BY_KEY) with the records as values) against the oracle's
composition and the reference's golden vectors.  Bar: bit-exact."""
import numpy as np
import pytest

import compute_amd as ca
import oracle as O
from test_oracle import sort_by_transform_oracle

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(99)


def run(q, rec, get, fn, descending):
    rec = np.ascontiguousarray(rec)
    width = 1 if rec.ndim == 1 else rec.shape[1]
    vdt = np.dtype("V%d" % (rec.itemsize * width)) if width > 1 else rec.dtype
    flat = rec.view(vdt).reshape(-1) if width > 1 else rec
    d = ca.DeviceArray.from_host(flat, q)
    ca.sort_by_transform(q, d, rec.dtype, 0 if get is None else get * rec.itemsize, fn, descending)
    return np.frombuffer(d.to_host(q).tobytes(), rec.dtype).reshape(rec.shape)


def test_golden_sort_by_transform(queue, golden):
    seen = 0
    for c in golden:
        if c["algo"] != "sort_by_transform":
            continue
        seen += 1
        got = run(queue, np.array(c["records"], c["dtype"]), c["get"], c["fn"], c["descending"])
        assert got.tolist() == c["expected"], c["id"]
    assert seen == 2


@pytest.mark.parametrize("dt,width", [("int32", 1), ("int32", 2), ("int32", 4), ("float32", 4), ("float32", 2),
                                      ("int64", 2), ("int16", 4), ("int8", 2), ("float64", 2), ("uint16", 1)])
@pytest.mark.parametrize("n", [2, 31, 32, 33, 1000, 70001])
def test_sort_by_transform_matches_oracle(queue, dt, width, n):
    info = np.iinfo(dt) if np.dtype(dt).kind in "iu" else None
    if info is not None:
        rec = RNG.integers(max(info.min, -1000), min(info.max, 1000), (n, width) if width > 1 else n).astype(dt)
        if info.min < 0 and n > 4:
            rec.flat[0] = info.min  # abs(INT_MIN) wraps to INT_MIN
    else:
        rec = RNG.integers(-100, 100, (n, width) if width > 1 else n).astype(dt) * np.array(0.5, dt)
    for get in ([None] if width == 1 else [0, width - 1]):
        for fn in ("id", "abs"):
            for desc in (False, True):
                got = run(queue, rec, get, fn, desc)
                want = sort_by_transform_oracle(rec, get, fn, desc)
                assert got.tobytes() == want.tobytes(), (dt, width, n, get, fn, desc)
