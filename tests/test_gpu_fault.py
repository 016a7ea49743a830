"""The fault path from a kernel to the host, forced from the host: the probe
runs the bounded wait every look-back uses on a word nothing sets, with a
spin limit of a few polls, so the wait gives up and records the fault while
every kernel exits normally.  bcmi_queue_finish must then report
BCMI_OUT_OF_RESOURCES (CL_OUT_OF_RESOURCES) with the give-up message, clear
the fault, and leave the queue fully usable (its spin limit restored)."""
import numpy as np
import pytest

import compute_amd as ca

pytestmark = pytest.mark.gpu


def test_give_up_is_reported_once_and_cleared(queue):
    L = ca.lib()
    assert L.bcmi_queue_finish(queue.h) == 0
    assert L.bcmi_debug_fault_probe(queue.h, 64) == 0
    assert L.bcmi_queue_finish(queue.h) == ca.OUT_OF_RESOURCES
    assert b"spin limit" in L.bcmi_last_error()
    assert L.bcmi_queue_finish(queue.h) == 0
    with pytest.raises(ca.BcmiError) as e:
        ca.check(L.bcmi_debug_fault_probe(queue.h, 1)) or ca.check(L.bcmi_queue_finish(queue.h))
    assert e.value.code == ca.OUT_OF_RESOURCES


def test_queue_works_after_a_reported_fault(queue):
    L = ca.lib()
    ca.check(L.bcmi_debug_fault_probe(queue.h, 8))
    assert L.bcmi_queue_finish(queue.h) == ca.OUT_OF_RESOURCES
    rng = np.random.default_rng(5)
    k = rng.integers(0, 2 ** 32, 3_000_017, dtype=np.uint64).astype(np.uint32)
    d = ca.DeviceArray.from_host(k, queue)
    ca.sort(queue, d, api=ca.RADIX_SORT)  # look-backs with the default spin limit again
    assert np.array_equal(d.to_host(queue), np.sort(k))
    assert L.bcmi_queue_finish(queue.h) == 0


def test_probe_rejects_bad_limits(queue):
    L = ca.lib()
    assert L.bcmi_debug_fault_probe(queue.h, 0) == ca.INVALID_VALUE
    assert L.bcmi_debug_fault_probe(queue.h, 1 << 20) == ca.INVALID_VALUE
    assert L.bcmi_queue_finish(queue.h) == 0
