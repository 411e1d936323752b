"""The exact fixed-point dipole sums of the slice histograms
(mtrack_amd/csrc/fixedpoint.h, shared by the GPU kernels and this host
driver): each double x becomes round(x 2^100) (ties away from zero) as a
128-bit two's-complement integer, checked against exact rational arithmetic;
sums of Np values below the per-particle bound the ring uses, min(2^20,
2^27 / Np), stay exact at the extremes (no wrap of the 128-bit range), while
the old bound of 2^20 alone would wrap beyond 128 particles; the conversion
back to double is within one ulp of the exact sum."""
import os
import struct
import subprocess
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TWO100 = 1 << 100


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fx") / "fx_check")
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    subprocess.run([hipcc, "-x", "hip", "-O2", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "mtrack_amd", "csrc"),
                    os.path.join(ROOT, "tests", "fx_check.cpp"), "-o", exe], check=True, capture_output=True)
    return exe


def run(exe, xs):
    bits = [struct.unpack("<Q", struct.pack("<d", float(x)))[0] for x in xs]
    inp = "%d\n" % len(xs) + "\n".join("%016x" % b for b in bits) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    q = []
    for line in out[:len(xs)]:
        hi, lo = (int(t, 16) for t in line.split())
        v = (hi << 64) | lo
        q.append(v - (1 << 128) if v >> 127 else v)
    _, shi, slo, vb = out[len(xs)].split()
    s = (int(shi, 16) << 64) | int(slo, 16)
    s = s - (1 << 128) if s >> 127 else s
    return q, s, struct.unpack("<d", struct.pack("<Q", int(vb, 16)))[0]


def exact_q(x):
    f = Fraction(float(x)) * TWO100
    n = abs(f)
    r = int(n) + (1 if n - int(n) >= Fraction(1, 2) else 0)
    return r if f >= 0 else -r


def check_sum(exe, xs):
    q, s, v = run(exe, xs)
    assert q == [exact_q(x) for x in xs]
    exact = sum(q)
    assert s == exact, "128-bit sum wrapped"
    ref = float(Fraction(exact, TWO100))
    assert abs(v - ref) <= np.spacing(abs(ref)) if ref != 0 else v == 0
    return exact


def test_split_and_sum_random(driver):
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.normal(0, 1e-3, 2000), rng.uniform(-1000, 1000, 500), [0.0, -0.0, 5e-324, 3e-31,
                                                                                   -3e-31, 7.9e-31, 1.2e-30]])
    check_sum(driver, xs)


@pytest.mark.parametrize("np_", [129, 1000, 100000])
def test_bound_keeps_sums_exact(driver, np_):
    lim = 134217728.0 / np_ if np_ > 128 else 1048576.0
    top = np.nextafter(lim, 0.0)
    for sign in (1.0, -1.0):
        exact = check_sum(driver, [sign * top] * np_)
        assert abs(exact) < (1 << 127)


def test_old_bound_wraps(driver):
    """129 particles just below 2^20 sum past 2^27: the 128-bit sum wraps."""
    q, s, _ = run(driver, [np.nextafter(1048576.0, 0.0)] * 129)
    assert sum(q) >= (1 << 127) and s < 0
