"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
the checker.  The oracle consumes the product's parsed configuration struct
(a data layout, include/mbtrack_mi355x.h) but runs its own restatement of the
turn (oracle.c).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REFHARNESS = os.path.join(REF_DIR, "refharness")
REF_MBTRACK = os.path.join(REF_DIR, "mbtrack-mpi")

RNG_REFERENCE = 0
RNG_PHILOX = 1
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = ctypes.CDLL(LIBPATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.orc_ring_create.argtypes = [vp, ip]
        L.orc_ring_create.restype = vp
        L.orc_ring_destroy.argtypes = [vp]
        L.orc_ring_set_bunch.argtypes = [vp, ip, dp]
        L.orc_ring_get_bunch.argtypes = [vp, ip, dp]
        L.orc_ring_prepare.argtypes = [vp, dp]
        L.orc_ring_track.argtypes = [vp, ctypes.c_long, dp, dp, dp]
        L.orc_ring_final_stats.argtypes = [vp, dp]
        L.orc_ring_histograms.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_ring_greens.argtypes = [vp, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_generate_bunch.argtypes = [vp, ip, dp]
        L.orc_glibc_rand_seq.argtypes = [ctypes.c_uint, ip, ctypes.POINTER(ctypes.c_int)]
        L.orc_philox4x32.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint),
                                     ctypes.POINTER(ctypes.c_uint)]
        L.orc_qe_normals.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_ulonglong, ctypes.c_uint, dp]
        L.orc_logf.argtypes = [ctypes.c_float]
        L.orc_logf.restype = ctypes.c_float
        L.orc_sincos2pif.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        if hasattr(L, "orc_wofz"):
            L.orc_wofz.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_ws_create.argtypes = [vp, ip, dp]
        L.orc_ws_create.restype = vp
        L.orc_ws_destroy.argtypes = [vp]
        L.orc_ws_track.argtypes = [vp, ctypes.c_long, dp, dp, dp, dp]
        L.orc_ws_get_slices.argtypes = [vp, ip, dp]
        L.orc_ws_population.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


class OracleRing:
    """CPU restatement of the ring; same call shape as mtrack_amd.Ring."""

    def __init__(self, cfg, rng=RNG_REFERENCE):
        self.cfg = cfg
        self.Nb = cfg.geti("Nbunch")
        self.Np = cfg.geti("Np")
        self.Nc = cfg.geti("Ncell")
        p = lib().orc_ring_create(cfg.ptr, rng)
        if not p:
            raise RuntimeError("oracle cannot represent this configuration")
        self.ptr = ctypes.c_void_p(p)

    def close(self):
        if getattr(self, "ptr", None):
            lib().orc_ring_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    def set_bunch(self, ib, soa6):
        a = np.ascontiguousarray(soa6, dtype=np.float64).reshape(-1)
        lib().orc_ring_set_bunch(self.ptr, ib, _dp(a))

    def get_bunch(self, ib):
        a = np.zeros(6 * self.Np)
        lib().orc_ring_get_bunch(self.ptr, ib, _dp(a))
        return a.reshape(6, self.Np)

    def prepare(self):
        st = np.zeros((self.Nb, 12))
        lib().orc_ring_prepare(self.ptr, _dp(st))
        return st

    def track(self, nturns):
        st = np.zeros((nturns, self.Nb, 12))
        am = np.zeros((nturns, self.Nb, 6))
        rf = np.zeros((nturns, 2))
        lib().orc_ring_track(self.ptr, nturns, _dp(st), _dp(am), _dp(rf))
        return st, am, rf

    def final_stats(self):
        st = np.zeros((self.Nb, 12))
        lib().orc_ring_final_stats(self.ptr, _dp(st))
        return st

    def greens(self):
        nc = max(self.Nc, 1)
        g = [np.zeros(nc) for _ in range(3)]
        planes = (ctypes.c_int * 3)()
        lib().orc_ring_greens(self.ptr, _dp(g[0]), _dp(g[1]), _dp(g[2]), planes)
        return g[0], g[1], g[2], list(planes)

    def histograms(self):
        c = np.zeros((self.Nb, max(self.Nc, 1)))
        t = (ctypes.c_int * (2 * self.Nb))()
        lib().orc_ring_histograms(self.ptr, _dp(c), t)
        return c, np.array(list(t)).reshape(self.Nb, 2)


WS_IBUNCH_REFERENCE = 0
WS_IBUNCH_FILL = 1


class OracleWS:
    """CPU restatement of the weak-strong turn with fast beam-ion interaction
    (oracle_ws.c); the same call shape as mtrack_amd.WeakStrong."""

    def __init__(self, cfg, ibunch_mode=WS_IBUNCH_FILL):
        self.cfg = cfg
        self.Nb = cfg.geti("Nbunch")
        self.h = cfg.geti("Nharm")
        self.S = 2 * cfg.geti("Nslc") + 1
        self.Np = cfg.geti("Np")
        self.dist = np.zeros(6 * self.Np)
        p = lib().orc_ws_create(cfg.ptr, ibunch_mode, _dp(self.dist))
        if not p:
            raise RuntimeError("oracle cannot represent this weak-strong configuration")
        self.ptr = ctypes.c_void_p(p)
        self.dist = self.dist.reshape(6, self.Np)

    def close(self):
        if getattr(self, "ptr", None):
            lib().orc_ws_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    def track(self, nturns):
        cm = np.zeros((nturns, self.Nb, 4))
        mon = np.zeros((nturns, self.Nb, 18))
        ion = np.zeros((nturns, self.Nb, 3))
        zcm = np.zeros((nturns, self.h))
        rc = lib().orc_ws_track(self.ptr, nturns, _dp(cm), _dp(mon), _dp(ion), _dp(zcm))
        return cm, mon, ion, zcm, rc

    def slices(self, ib):
        out = np.zeros(6 * self.S)
        lib().orc_ws_get_slices(self.ptr, ib, _dp(out))
        return out.reshape(6, self.S)

    def population(self):
        npop = (ctypes.c_int * self.S)()
        nofe = (ctypes.c_int * (25 * self.S))()
        lib().orc_ws_population(self.ptr, npop, nofe)
        return np.array(list(npop)), np.array(list(nofe)).reshape(self.S, 5, 5)


def wofz(x, y):
    u, v = ctypes.c_double(), ctypes.c_double()
    f = lib().orc_wofz(float(x), float(y), ctypes.byref(u), ctypes.byref(v))
    return u.value, v.value, f


def generate_bunch(cfg, kb):
    n = cfg.geti("Np")
    out = np.zeros(6 * n)
    if lib().orc_generate_bunch(cfg.ptr, kb, _dp(out)) != 0:
        raise RuntimeError("oracle generator does not cover this distribution")
    return out.reshape(6, n)


def glibc_rand_seq(seed, n):
    out = (ctypes.c_int * n)()
    lib().orc_glibc_rand_seq(seed & 0xFFFFFFFF, n, out)
    return np.array(list(out), dtype=np.int64)


def philox(ctr, key):
    c = (ctypes.c_uint * 4)(*ctr)
    k = (ctypes.c_uint * 2)(*key)
    o = (ctypes.c_uint * 4)()
    lib().orc_philox4x32(c, k, o)
    return list(o)


def qe_normals(seed, bucket, rev, j):
    out = np.zeros(5)
    lib().orc_qe_normals(seed, bucket, rev, j, _dp(out))
    return out
