"""Faddeeva function.

The oracle's restatement of ACM TOMS 680 is pinned by the reference's own known
answers (src/test_complexerrf.c:33-59, tolerance as there), by scipy.special.wofz
(an independent implementation, Faddeeva package) to 2e-13 relative over the
plane, and by the symmetries w(-z) = 2 exp(-z^2) - w(z), w(conj z) = conj w(-z).
The GPU kernel (mbt_wofz) must match the oracle to 1e-13 relative in FP64 and
scipy to 2e-5 relative in FP32.
"""
import ctypes

import numpy as np
import pytest
from scipy.special import wofz

import mtrack_amd as m
import oracle as o

KNOWN = [((0.0, 0.0), (1.0, 0.0), 1e-8), ((0.01, 0.01), (0.9887176929549547, 0.01108529605747726), 1e-8),
         ((0.5, 0.5), (0.5331567079121750, 0.2304882313844584), 1e-8), ((0.5, 1.0), (0.391234, 0.127202), 1e-5),
         ((1.0, 1.0), (0.3047442052569128, 0.2082189382028316), 1e-8),
         ((2.5, 2.5), (0.1167371250446503, 0.1079085859964814), 1e-8),
         ((15.0, 15.0), (0.01882714532513676, 0.01878535427799565), 1e-8)]


def orc(x, y):
    u, v = ctypes.c_double(), ctypes.c_double()
    f = o.lib().orc_wofz(x, y, ctypes.byref(u), ctypes.byref(v))
    return f, complex(u.value, v.value)


def grid(n=4000, seed=3):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.uniform(-25, 25, n), np.linspace(-6, 6, 61).repeat(61)])
    y = np.concatenate([rng.uniform(-25, 25, n), np.tile(np.linspace(-6, 6, 61), 61)])
    return x, y


@pytest.mark.parametrize("z,w,tol", KNOWN)
def test_known_answers(z, w, tol):
    f, val = orc(*z)
    assert f == 0
    assert abs(val.real - w[0]) <= tol and abs(val.imag - w[1]) <= tol


def test_against_scipy_and_symmetry():
    x, y = grid()
    worst = 0.0
    for a, b in zip(x, y):
        f, val = orc(a, b)
        if f:
            continue
        ref = wofz(complex(a, b))
        worst = max(worst, abs(val - ref) / abs(ref))
    assert worst < 2e-13
    for a, b in [(0.3, 0.7), (2.0, 0.1), (5.0, 3.0), (0.05, 0.02)]:
        z = complex(a, b)
        _, wz = orc(a, b)
        _, wmz = orc(-a, -b)
        assert abs(wmz - (2 * np.exp(-z * z) - wz)) <= 1e-12 * abs(wmz)
        _, wc = orc(a, -b)
        assert abs(wc - np.conj(wmz)) <= 1e-13 * abs(wc)


def test_overflow_flag():
    f, _ = orc(1.0, -30.0)
    assert f == 1


@pytest.mark.gpu
def test_gpu_wofz_fp64_fp32():
    x, y = grid(20000, 7)
    n = len(x)
    L = m.lib()
    wr, wi, fl = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
    rc = L.mbt_wofz(0, n, m.dptr(x), m.dptr(y), m.dptr(wr), m.dptr(wi),
                    fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert rc == 0
    worst = 0.0
    for i in range(n):
        f, val = orc(x[i], y[i])
        assert f == fl[i]
        if f:
            continue
        worst = max(worst, abs(complex(wr[i], wi[i]) - val) / abs(val))
    print("fp64 gpu vs oracle", worst)
    assert worst < 1e-13
    xf, yf = x.astype(np.float32), y.astype(np.float32)
    wrf, wif = np.zeros(n, np.float32), np.zeros(n, np.float32)
    L.mbt_wofz_f32.argtypes = [ctypes.c_int, ctypes.c_long] + [ctypes.POINTER(ctypes.c_float)] * 4 + \
        [ctypes.POINTER(ctypes.c_int)]
    fp = ctypes.POINTER(ctypes.c_float)
    rc = L.mbt_wofz_f32(0, n, xf.ctypes.data_as(fp), yf.ctypes.data_as(fp), wrf.ctypes.data_as(fp),
                        wif.ctypes.data_as(fp), fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert rc == 0
    ok = (fl == 0) & (np.abs(yf) < 8) & (np.abs(xf) < 8) | ((fl == 0) & (yf >= 0))
    ref = wofz(xf[ok].astype(np.float64) + 1j * yf[ok].astype(np.float64))
    err = np.abs(wrf[ok] + 1j * wif[ok] - ref) / np.abs(ref)
    print("fp32 gpu vs scipy", err.max())
    assert err.max() < 2e-5
