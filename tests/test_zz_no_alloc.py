"""The allocation-free Faddeeva entry point (mbt_wofz_noalloc: staging in a
fixed device array, no device allocation) against the allocating one
(mbt_wofz) through the C ABI: bit-identical values and overflow flags on two
calls in a row, the second with fewer points (the array keeps the first call's
contents beyond them), and a clear error for more points than the array holds."""
import ctypes

import numpy as np
import pytest

import mtrack_amd as m

pytestmark = pytest.mark.gpu
NOALLOC_MAX = 1 << 18


def _call(fn, x, y):
    n = len(x)
    wr, wi, fl = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
    rc = fn(0, n, m.dptr(x), m.dptr(y), m.dptr(wr), m.dptr(wi), fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return rc, wr, wi, fl


def test_noalloc_equals_allocating_entry_point():
    L = m.lib()
    rng = np.random.default_rng(20241203)
    for n in (NOALLOC_MAX, 3001):
        print("case mbt_wofz_noalloc, %d points" % n, flush=True)
        # all four quadrants, every branch of the algorithm, and points whose
        # result overflows (flag 1)
        x = rng.uniform(-15.0, 15.0, n)
        y = rng.uniform(-30.0, 15.0, n)
        rc_a, wr_a, wi_a, fl_a = _call(L.mbt_wofz, x, y)
        rc_b, wr_b, wi_b, fl_b = _call(L.mbt_wofz_noalloc, x, y)
        assert rc_a == 0 and rc_b == 0, m.lib().mbt_last_error().decode()
        assert np.any(fl_a == 1) and np.any(fl_a == 0)
        assert np.array_equal(fl_a, fl_b)
        assert np.array_equal(wr_a.view(np.uint64), wr_b.view(np.uint64))
        assert np.array_equal(wi_a.view(np.uint64), wi_b.view(np.uint64))


def test_noalloc_rejects_more_points_than_its_array():
    L = m.lib()
    big = np.zeros(NOALLOC_MAX + 1)
    rc, _, _, _ = _call(L.mbt_wofz_noalloc, big, big)
    assert rc == m.MBT_EINVAL
    assert "scratch" in L.mbt_last_error().decode()
