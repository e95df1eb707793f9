"""The allocation-free exclusive scan (sptrsv_exclusive_scan_noalloc) against
the allocating one (sptrsv_exclusive_scan) and the oracle, through the C ABI.

Bit-exact (integer work).  Two calls in a row on the shared static scratch,
the second with a smaller input, so stale tile totals of the first call would
show; lengths that use one, two and three levels of tile totals; a length
beyond the stated maximum is rejected without running anything; and the
entry point runs inside a hipGraph capture (which an allocation would break),
replayed with eager work between replays.
"""
import ctypes

import numpy as np
import pytest

import sptrsv_amd as S
import sptrsv_oracle as orc
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 28


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def _scan_noalloc(a, stream=None):
    d = S.DeviceArray.from_host(np.ascontiguousarray(a, np.int32))
    st = S.lib().sptrsv_exclusive_scan_noalloc(d.ptr, len(a), stream.ptr if stream else None)
    assert st == 0, S.STATUS.get(st, st)
    S.check(S.lib().sptrsv_synchronize(), "sync")
    return d.to_host()


def _scan_alloc(a):
    return S.exclusive_scan(np.ascontiguousarray(a, np.int32))


def test_two_calls_second_smaller_match_allocating_scan():
    rng = np.random.default_rng(7)
    big = rng.integers(0, 9, size=9_000_001).astype(np.int32)   # three levels
    small = rng.integers(0, 9, size=70_001).astype(np.int32)     # two levels
    out_big = _scan_noalloc(big)
    np.testing.assert_array_equal(out_big, _scan_alloc(big))
    np.testing.assert_array_equal(out_big, orc.exclusive_scan(big))
    out_small = _scan_noalloc(small)
    np.testing.assert_array_equal(out_small, _scan_alloc(small))
    np.testing.assert_array_equal(out_small, orc.exclusive_scan(small))


@pytest.mark.parametrize("length", [1, 2, 2047, 2048, 2049, 4_194_304, 4_194_305])
def test_lengths_across_levels(length):
    a = np.random.default_rng(length).integers(0, 5, size=length).astype(np.int32)
    np.testing.assert_array_equal(_scan_noalloc(a), _scan_alloc(a))


def test_rejects_lengths_beyond_the_maximum():
    d = S.DeviceArray((16,), np.int32)
    d.upload(np.arange(16, dtype=np.int32))
    assert S.lib().sptrsv_exclusive_scan_noalloc(d.ptr, NOALLOC_MAX + 1, None) == -2
    assert S.lib().sptrsv_exclusive_scan_noalloc(d.ptr, -1, None) == -2
    assert S.lib().sptrsv_exclusive_scan_noalloc(None, 5, None) == -2
    np.testing.assert_array_equal(d.to_host(), np.arange(16))  # untouched
    assert S.lib().sptrsv_exclusive_scan_noalloc(d.ptr, 0, None) == 0


def test_runs_under_graph_capture():
    # hipStreamBeginCapture / EndCapture through the HIP runtime the library
    # links; the captured scan is replayed with eager scans in between
    hip = None
    for name in ("libamdhip64.so.7", "libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
            break
        except OSError:
            pass
    assert hip is not None
    a = np.random.default_rng(3).integers(0, 7, size=300_000).astype(np.int32)
    ref = orc.exclusive_scan(a)
    # the first call on a device looks the scratch address up; capture after it
    np.testing.assert_array_equal(_scan_noalloc(a), ref)
    d = S.DeviceArray.from_host(a)
    stream = S.Stream()
    assert hip.hipStreamBeginCapture(stream.ptr, ctypes.c_int(0)) == 0
    st = S.lib().sptrsv_exclusive_scan_noalloc(d.ptr, len(a), stream.ptr)
    graph = ctypes.c_void_p()
    assert hip.hipStreamEndCapture(stream.ptr, ctypes.byref(graph)) == 0
    assert st == 0
    exe = ctypes.c_void_p()
    assert hip.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, ctypes.c_size_t(0)) == 0
    other = np.random.default_rng(4).integers(0, 7, size=50_000).astype(np.int32)
    try:
        for _ in range(3):
            d.upload(a)
            assert hip.hipGraphLaunch(exe, stream.ptr) == 0
            stream.synchronize()
            np.testing.assert_array_equal(d.to_host(), ref)
            np.testing.assert_array_equal(_scan_noalloc(other), orc.exclusive_scan(other))
    finally:
        hip.hipGraphExecDestroy(exe)
        hip.hipGraphDestroy(graph)
