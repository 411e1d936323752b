"""How far rounding-level changes move the weak-strong outputs: the floor
under any comparison between the oracle and an implementation that rounds
differently (tests/test_gpu_ws.py's bound rests on it).

Variants of the oracle (oracle/oracle_sens.h): grid quotients by the
reciprocal spacing, sums over the ion cloud as pairwise trees, and sin / cos /
exp moved by one ulp -- each a change the GPU makes or may make.  On the
416-bucket configuration each one alone moves some output (the cloud's Nri y
sum first, a sum of terms of both signs) by 1e-11 to 2e-11 of its largest
magnitude over the run, while the GPU differs from the oracle by less (measured
6.2e-12).  The test pins that floor: every variant stays within WS_FLOOR_MAX
of the oracle, and at least one reaches WS_FLOOR_MIN, so the GPU bound is not
below what rounding alone produces."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from variants import write_ws_variant

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")
VARIANTS = ("recip", "tree", "ulp")
WS_FLOOR_MIN = 5e-12
WS_FLOOR_MAX = 5e-11


def _load(name):
    path = os.path.join(ORACLE_DIR, "liboracle_sens_%s.so" % name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "sens"], check=True)
    L = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.orc_ws_create.argtypes = [vp, ip, dp]
    L.orc_ws_create.restype = vp
    L.orc_ws_destroy.argtypes = [vp]
    L.orc_ws_track.argtypes = [vp, ctypes.c_long, dp, dp, dp, dp]
    L.orc_ws_get_slices.argtypes = [vp, ip, dp]
    return L


def _run(L, cfg, n, mode):
    Np, Nb, h = cfg.geti("Np"), cfg.geti("Nbunch"), cfg.geti("Nharm")
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    dist = np.zeros(6 * Np)
    w = L.orc_ws_create(cfg.ptr, mode, dp(dist))
    out = [np.zeros((n, Nb, 4)), np.zeros((n, Nb, 18)), np.zeros((n, Nb, 3)), np.zeros((n, h))]
    L.orc_ws_track(w, n, *[dp(a) for a in out])
    S = 2 * cfg.geti("Nslc") + 1
    sl = np.zeros((Nb, 6 * S))
    for ib in range(Nb):
        t = np.zeros(6 * S)
        L.orc_ws_get_slices(w, ib, dp(t))
        sl[ib] = t
    L.orc_ws_destroy(w)
    return out + [sl.reshape(Nb, 6, S)]


def ws_errors(g, r, nrev_mon):
    """Per-quantity errors as tests/test_gpu_ws.py takes them: max |g - r| over
    the run relative to max |r| of the same quantity."""
    from test_gpu_ws import scaled_err
    (gcm, gmon, gion, gzcm, gsl), (ocm, omon, oion, ozcm, osl) = g, r
    n = len(ocm)
    e = {"cm": scaled_err(gcm, ocm), "zcm": scaled_err(gzcm, ozcm)}
    mon_t = [t for t in range(n) if t % nrev_mon == 0]
    for k, lab in enumerate(["pos", "pos_sigma", "slope", "slope_sigma", "ampinv", "ampinv_cm"]):
        e[lab] = scaled_err(gmon[mon_t][..., 3 * k:3 * k + 3], omon[mon_t][..., 3 * k:3 * k + 3])
    for k, lab in enumerate(["ion_x", "ion_y", "ion_q"]):
        e[lab] = scaled_err(gion[..., k], oion[..., k]) if np.any(oion[..., k]) else float(np.max(np.abs(gion[..., k])))
    for k, lab in enumerate(["tau", "delta", "x", "xp", "z", "zp"]):
        e["slice_" + lab] = scaled_err(gsl[:, k], osl[:, k]) if np.any(osl[:, k]) else float(np.max(np.abs(gsl[:, k])))
    return e


def test_ws_rounding_floor(tmp_path):
    path, mode = write_ws_variant("ws_h416", str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    n = cfg.geti("NrevTot")
    orc = o.OracleWS(cfg, mode)
    base = list(orc.track(n)[:4]) + [np.stack([orc.slices(ib) for ib in range(orc.Nb)])]
    orc.close()
    worst = {}
    for v in VARIANTS:
        e = ws_errors(_run(_load(v), cfg, n, mode), base, cfg.geti("NrevMon"))
        worst[v] = max(e.values())
        top = sorted(e.items(), key=lambda kv: -kv[1])[:4]
        print("%s: largest output change %.3g (%s)" % (v, worst[v], ", ".join("%s %.2g" % kv for kv in top)))
    assert max(worst.values()) <= WS_FLOOR_MAX
    assert max(worst.values()) >= WS_FLOOR_MIN
