"""Random streams.

* glibc rand(): both restatements (oracle and product) equal the C library's own
  rand() for seeds including 0 and >= 2^31 (the reference draws its initial
  distributions and its quantum-excitation normals from it, src/bunch.c:379-436).
* Philox4x32-10: the published known-answer vectors of the Random123 suite.
* QE normals (counter-based, float Box-Muller from explicit-FMA polynomials):
  mean, variance and tail fractions of N(0,1); the log and sin/cos kernels
  against libm.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as o


@pytest.mark.parametrize("seed", [0, 1, 2, 123456, 2**31 - 1, 2**31, 2**32 - 1, 4000000000])
def test_glibc_rand_restatement(seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    expect = [libc.rand() for _ in range(2000)]
    assert list(o.glibc_rand_seq(seed, 2000)) == expect


def test_philox_known_answers():
    assert o.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert o.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert o.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_qe_normals_distribution():
    n = np.array([o.qe_normals(1001, 3, rev, j) for rev in range(40) for j in range(1000)]).reshape(-1)
    assert abs(n.mean()) < 5 / math.sqrt(n.size)
    assert abs(n.var() - 1) < 5 * math.sqrt(2 / n.size)
    frac = np.mean(np.abs(n) > 2)
    assert abs(frac - 0.0455) < 0.004


def test_qe_kernels_accuracy():
    L = o.lib()
    for u in np.concatenate([np.linspace(1e-6, 1, 2001), [5.96e-8, 1.0]]):
        v = L.orc_logf(ctypes.c_float(u))
        assert abs(v - math.log(np.float32(u))) <= 3e-7 * max(1, abs(math.log(u)))
    s, c = ctypes.c_float(), ctypes.c_float()
    for t in np.linspace(0, 1, 4001):
        L.orc_sincos2pif(ctypes.c_float(t), ctypes.byref(s), ctypes.byref(c))
        tt = float(np.float32(t))
        assert abs(s.value - math.sin(2 * math.pi * tt)) < 2e-6
        assert abs(c.value - math.cos(2 * math.pi * tt)) < 2e-6
