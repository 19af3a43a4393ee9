"""es_engine_scan_window_static (the window scan that allocates nothing) against the engine's
own es_engine_scan_window + es_engine_collect: two calls in a row, the second on a smaller
window, through the C ABI.  Device buffers for the caller come from the HIP runtime the library
itself links (libamdhip64), so the test allocates them and the entry point never does."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import es_native as nat
import es_plan as EP

pytestmark = pytest.mark.gpu

C1_WHEEL = [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]]


def _hip():
    nat.lib()                                    # loads libamdhip64 with the engine
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _static_scan(hip, eng, lo, hi, d_ids, d_res, cap):
    L = nat.lib()
    nat.check(L.es_engine_scan_window_static(eng._h, lo, hi, d_ids, cap, d_res, None))
    assert hip.hipDeviceSynchronize() == 0
    res = np.zeros(4, dtype=np.uint64)
    assert hip.hipMemcpy(res.ctypes.data, d_res, 32, 2) == 0          # device to host
    n = int(res[0])
    ids = np.zeros(max(1, min(n, cap)), dtype=np.uint64)
    if n:
        assert hip.hipMemcpy(ids.ctypes.data, d_ids, 8 * min(n, cap), 2) == 0
    mn = (int(res[2]) << 64 | int(res[1])) if res[3] else None
    return n, np.sort(ids[:min(n, cap)]), mn


def test_static_scan_equals_allocating_scan_twice():
    import gpu_sieve_mi355x as G
    hip = _hip()
    plan = EP.Plan(30, {"max_survivors": 1 << 20}, C1_WHEEL)
    eng = G.GpuSieve(plan)
    cap = 65536
    d_ids, d_res = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_ids), 8 * cap) == 0
    assert hip.hipMalloc(C.byref(d_res), 32) == 0
    try:
        for t, lo, hi in ((3, 0, 1 << 19), (3, 5_000_000, 5_000_000 + (1 << 16)),
                          (0, 13_200_000, 13_300_000)):
            print(f"case static scan k=30 t={t} window=[{lo},{hi})", flush=True)
            want_ids, want_min = eng.scan_window(t, lo, hi)      # sets the block, allocating path
            n, got, mn = _static_scan(hip, eng, lo, hi, d_ids, d_res, cap)
            assert n == len(want_ids) and n <= cap
            assert np.array_equal(got, np.sort(np.asarray(want_ids, dtype=np.uint64)))
            got_min = None if mn is None else t * plan.M + mn
            assert got_min == want_min
        # cap beyond the scratch is refused before any launch
        rc = nat.lib().es_engine_scan_window_static(eng._h, 0, 1 << 10, d_ids, cap + 1, d_res,
                                                    None)
        assert rc == nat.ES_ERR_LIMIT and b"65536" in nat.lib().es_last_error()
    finally:
        hip.hipFree(d_ids)
        hip.hipFree(d_res)


K399_PRIMARY = [[107, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1], [137, 1],
                [41, 1], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1], [31, 1], [67, 1], [5, 2]]


def test_static_scan_row_lane_engine():
    """The k=399 primary wheel's engine (row-lane, the kernel instance sized for 8 waves per
    SIMD, as the bench runs it): the static scan runs the spill-free 6-wave instance and returns
    the same survivors and minimum on windows of block 6 around g(399), twice, the second window
    smaller; the minimum is g(399) itself."""
    import gpu_sieve_mi355x as G
    from conftest import bfile
    hip = _hip()
    plan = EP.Plan(399, {"max_survivors": 1 << 20}, K399_PRIMARY)
    eng = G.GpuSieve(plan)
    assert eng.layout == "rows" and eng.threads == 1024
    g = bfile()[399]
    gi, gt = plan.n_to_idx(g)
    assert gt == 6
    cap = 65536
    d_ids, d_res = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_ids), 8 * cap) == 0
    assert hip.hipMalloc(C.byref(d_res), 32) == 0
    try:
        for half in (1 << 31, 1 << 20):
            a, b = gi - half, gi + half
            print(f"case static scan k=399 rows threads={eng.threads} t=6 window=[{a},{b})",
                  flush=True)
            want_ids, want_min = eng.scan_window(6, a, b)
            n, got, mn = _static_scan(hip, eng, a, b, d_ids, d_res, cap)
            assert n >= 1 and n == len(want_ids)
            assert np.array_equal(got, np.sort(np.asarray(want_ids, dtype=np.uint64)))
            assert 6 * plan.M + mn == want_min == g
    finally:
        hip.hipFree(d_ids)
        hip.hipFree(d_res)
