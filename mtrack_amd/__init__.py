"""mtrack_amd: MI355X multi-bunch tracking turn behind a C ABI.

The product is the shared library libmbtrack_mi355x.so (host C + gfx950 HIP
kernels) declared in include/mbtrack_mi355x.h.  This module only binds it with
ctypes for tests and benchmarks; it has no compute path of its own, and
loading fails loudly when the library has not been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("MBTRACK_LIB") or os.path.join(HERE, "libmbtrack_mi355x.so")
EXE = os.path.join(HERE, "mbtrack-mi355x")

NSTAT = 12
NAMP = 6
MBT_EINVAL = -1
MBT_RING_NO_INIT_DIST = 1
MBT_RING_NO_GRAPHS = 2
MBT_RING_NO_RW_GEMM = 4
MBT_RING_MIXED = 8
MBT_RING_RW_OVERLAP = 16
MBT_RING_RW_GEMM = 32
MBT_RING_HOST_INIT_DIST = 64
MBT_WS_BUNCH_CURRENTS = 1
MBT_WS_SERIAL_TRAIN = 2
MBT_WS_MIXED = 4

_lib = None


class MbtError(RuntimeError):
    pass


def lib():
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise MbtError("libmbtrack_mi355x.so is not built: run python -m mtrack_amd.build")
        L = ctypes.CDLL(LIBPATH)
        vp, cp, ip, dp = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.mbt_last_error.restype = cp
        L.mbt_config_alloc.restype = vp
        L.mbt_config_release.argtypes = [vp]
        L.mbt_config_read.argtypes = [cp, cp, vp]
        L.mbt_config_get.argtypes = [vp, cp, dp]
        L.mbt_config_set.argtypes = [vp, cp, ctypes.c_double]
        L.mbt_config_bucket.argtypes = [vp, ip]
        L.mbt_config_dump.argtypes = [vp, cp]
        L.mbt_generate_bunch.argtypes = [vp, ip, dp]
        L.mbt_greens_functions.argtypes = [vp, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        L.mbt_ring_create.argtypes = [vp, ip, ip, ctypes.POINTER(vp)]
        L.mbt_ring_destroy.argtypes = [vp]
        L.mbt_ring_set_bunch.argtypes = [vp, ip, dp]
        L.mbt_ring_get_bunch.argtypes = [vp, ip, dp]
        L.mbt_ring_prepare.argtypes = [vp, dp]
        L.mbt_ring_track.argtypes = [vp, ctypes.c_long, dp, dp, dp]
        L.mbt_ring_async_max.argtypes = [vp]
        L.mbt_ring_async_max.restype = ctypes.c_long
        L.mbt_ring_track_async.argtypes = [vp, ctypes.c_long]
        L.mbt_ring_track_wait.argtypes = [vp, dp, dp, dp]
        L.mbt_ring_potentials.argtypes = [vp, ip, dp, dp]
        L.mbt_ring_final_stats.argtypes = [vp, dp]
        L.mbt_ring_histograms.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int)]
        L.mbt_ring_kernel_times.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_long)]
        L.mbt_ring_bytes_per_turn.argtypes = [vp]
        L.mbt_ring_bytes_per_turn.restype = ctypes.c_double
        L.mbt_ring_set_timing.argtypes = [vp, ip]
        L.mbt_ring_turn.argtypes = [vp]
        L.mbt_ring_turn.restype = ctypes.c_long
        L.mbt_ensemble_create.argtypes = [ctypes.POINTER(vp), ip, ip, ip, ctypes.POINTER(vp)]
        L.mbt_ensemble_set_bunch.argtypes = [vp, ip, dp]
        L.mbt_ensemble_get_bunch.argtypes = [vp, ip, dp]
        L.mbt_ensemble_size.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ctypes.c_longlong)]
        L.mbt_ring_set_chunk.argtypes = [vp, ip]
        L.mbt_ring_get_chunk.argtypes = [vp]
        L.mbt_ring_autotune.argtypes = [vp, ctypes.POINTER(ip), dp]
        L.mbt_ring_stream_ceiling.argtypes = [vp, ctypes.c_int, dp]
        L.mbt_ring_rf_state.argtypes = [vp, ip, dp]
        L.mbt_ring_graph_builds.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
        L.mbt_ring_placement.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp]
        L.mbt_ws_create.argtypes = [vp, ip, ip, ctypes.POINTER(vp)]
        L.mbt_ws_destroy.argtypes = [vp]
        L.mbt_ws_track.argtypes = [vp, ctypes.c_long, dp, dp, dp, dp]
        L.mbt_ws_async_max.argtypes = [vp]
        L.mbt_ws_async_max.restype = ctypes.c_long
        L.mbt_ws_track_async.argtypes = [vp, ctypes.c_long]
        L.mbt_ws_track_wait.argtypes = [vp, dp, dp, dp, dp]
        L.mbt_ws_get_slices.argtypes = [vp, ip, dp]
        L.mbt_ws_population.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip)]
        L.mbt_ws_distribution.argtypes = [vp, dp]
        L.mbt_ws_generate.argtypes = [vp, dp, ctypes.POINTER(ip), ctypes.POINTER(ip), dp]
        L.mbt_manager_weakstrong.argtypes = [vp, ip, ip]
        L.mbt_ws_set_timing.argtypes = [vp, ip]
        L.mbt_ws_train_time.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_long)]
        if hasattr(L, "mbt_manager_weakweak"):
            L.mbt_manager_weakweak.argtypes = [vp, ip]
        if hasattr(L, "mbt_wofz"):
            L.mbt_wofz.argtypes = [ip, ctypes.c_long, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        if hasattr(L, "mbt_wofz_noalloc"):
            L.mbt_wofz_noalloc.argtypes = [ip, ctypes.c_long, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MbtError("mbtrack error %d: %s" % (rc, lib().mbt_last_error().decode()))


def _flags(init_dist, graphs, rw_gemm=True, mixed=False, rw_overlap=False, force_gemm=False, host_init=False):
    return ((0 if init_dist else MBT_RING_NO_INIT_DIST) | (0 if graphs else MBT_RING_NO_GRAPHS)
            | (0 if rw_gemm else MBT_RING_NO_RW_GEMM) | (MBT_RING_MIXED if mixed else 0)
            | (MBT_RING_RW_OVERLAP if rw_overlap else 0) | (MBT_RING_RW_GEMM if force_gemm else 0)
            | (MBT_RING_HOST_INIT_DIST if host_init else 0))


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


class Config:
    """A parsed .conf file with its derived machine parameters."""

    def __init__(self, conf_path, work_path="work/"):
        L = lib()
        self.ptr = ctypes.c_void_p(L.mbt_config_alloc())
        self.path = conf_path
        _check(L.mbt_config_read(conf_path.encode(), work_path.encode(), self.ptr))

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.mbt_config_release(self.ptr)
            self.ptr = None

    def get(self, name):
        v = ctypes.c_double()
        _check(lib().mbt_config_get(self.ptr, name.encode(), ctypes.byref(v)))
        return v.value

    def set(self, name, value):
        _check(lib().mbt_config_set(self.ptr, name.encode(), float(value)))

    def geti(self, name):
        return int(self.get(name))

    @property
    def buckets(self):
        return [lib().mbt_config_bucket(self.ptr, k) for k in range(self.geti("Nbunch"))]

    def generate_bunch(self, kb):
        n = self.geti("Np")
        out = np.zeros(6 * n)
        _check(lib().mbt_generate_bunch(self.ptr, kb, dptr(out)))
        return out.reshape(6, n)

    def greens(self):
        nc = max(self.geti("Ncell"), 1)
        g = [np.zeros(nc) for _ in range(3)]
        planes = (ctypes.c_int * 3)()
        _check(lib().mbt_greens_functions(self.ptr, dptr(g[0]), dptr(g[1]), dptr(g[2]), planes))
        return g[0], g[1], g[2], list(planes)

    def dump(self, path):
        _check(lib().mbt_config_dump(self.ptr, path.encode()))


class Ring:
    """All bunches of one ring resident on one GPU (mbt_ring_* C ABI)."""

    def __init__(self, cfg, device=0, init_dist=True, graphs=True, mixed=False, rw_overlap=False, force_gemm=False,
                 host_init=False):
        L = lib()
        self.cfg = cfg
        self.Nb = cfg.geti("Nbunch")
        self.Np = cfg.geti("Np")
        self.Nc = cfg.geti("Ncell")
        p = ctypes.c_void_p()
        _check(L.mbt_ring_create(cfg.ptr, device, _flags(init_dist, graphs, mixed=mixed, rw_overlap=rw_overlap,
                                                         force_gemm=force_gemm, host_init=host_init),
                                 ctypes.byref(p)))
        self.ptr = p

    def close(self):
        if getattr(self, "ptr", None):
            lib().mbt_ring_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    def set_bunch(self, ib, soa6):
        """Coordinates [6, Np] of bunch ib, taken as the bunch at the end of a
        turn (any resistive-wall kick still pending for it is dropped)."""
        a = np.ascontiguousarray(soa6, dtype=np.float64).reshape(-1)
        if a.size != 6 * self.Np:
            raise MbtError("bunch %d holds %d macroparticles" % (ib, self.Np))
        _check(lib().mbt_ring_set_bunch(self.ptr, ib, dptr(a)))

    def get_bunch(self, ib):
        """Coordinates [6, Np] of bunch ib at the end of the last turn (the
        resistive-wall kick the next pass would apply is included)."""
        a = np.zeros(6 * self.Np)
        _check(lib().mbt_ring_get_bunch(self.ptr, ib, dptr(a)))
        return a.reshape(6, self.Np)

    def prepare(self):
        st = np.zeros((self.Nb, NSTAT))
        _check(lib().mbt_ring_prepare(self.ptr, dptr(st)))
        return st

    def track(self, nturns, want_stats=True, want_amp=True, want_rf=True):
        st = np.zeros((nturns, self.Nb, NSTAT)) if want_stats else None
        am = np.zeros((nturns, self.Nb, NAMP)) if want_amp else None
        rf = np.zeros((nturns, 2)) if want_rf else None
        _check(lib().mbt_ring_track(self.ptr, nturns, dptr(st), dptr(am), dptr(rf)))
        return st, am, rf

    def async_max(self):
        return lib().mbt_ring_async_max(self.ptr)

    def track_async(self, nturns):
        """Enqueues nturns turns (at most async_max()) without waiting."""
        _check(lib().mbt_ring_track_async(self.ptr, nturns))
        self._async_n = nturns

    def track_wait(self, want_stats=True, want_amp=True, want_rf=True):
        """Waits for the turns of track_async; returns their outputs as track does."""
        n = getattr(self, "_async_n", 0)
        st = np.zeros((n, self.Nb, NSTAT)) if want_stats else None
        am = np.zeros((n, self.Nb, NAMP)) if want_amp else None
        rf = np.zeros((n, 2)) if want_rf else None
        _check(lib().mbt_ring_track_wait(self.ptr, dptr(st), dptr(am), dptr(rf)))
        self._async_n = 0
        return st, am, rf

    def potentials(self, ib):
        """The tables of bunch ordinal ib the next turn's kick uses: 7 rows of Nc
        (wake L, LR wake, counts, wake V, dipoles V, wake H, dipoles H) and the
        previous phase, voltage and scan value."""
        out = np.zeros((7, max(self.Nc, 1)))
        prev = np.zeros(3)
        _check(lib().mbt_ring_potentials(self.ptr, ib, dptr(out), dptr(prev)))
        return out, prev

    def final_stats(self):
        st = np.zeros((self.Nb, NSTAT))
        _check(lib().mbt_ring_final_stats(self.ptr, dptr(st)))
        return st

    def histograms(self):
        c = np.zeros((self.Nb, max(self.Nc, 1)))
        t = (ctypes.c_int * (2 * self.Nb))()
        _check(lib().mbt_ring_histograms(self.ptr, dptr(c), t))
        return c, np.array(list(t)).reshape(self.Nb, 2)

    def set_timing(self, on=True):
        _check(lib().mbt_ring_set_timing(self.ptr, 1 if on else 0))

    def kernel_times(self):
        ms = np.zeros(4)
        n = (ctypes.c_long * 4)()
        _check(lib().mbt_ring_kernel_times(self.ptr, dptr(ms), n))
        return ms, list(n)

    def bytes_per_turn(self):
        return lib().mbt_ring_bytes_per_turn(self.ptr)

    def set_chunk(self, items):
        _check(lib().mbt_ring_set_chunk(self.ptr, items))

    def chunk(self):
        return lib().mbt_ring_get_chunk(self.ptr)

    def rf_state(self, ring=0):
        """(phai0, Vrf0 [MV], feedback on) of ring `ring` after the last turn."""
        out = np.zeros(3)
        _check(lib().mbt_ring_rf_state(self.ptr, ring, dptr(out)))
        return out[0], out[1], bool(out[2])

    def autotune(self):
        """Chooses the per-work-group chunk by timing; returns (items, {items: ms})."""
        best = ctypes.c_int()
        ms = np.zeros(7)
        _check(lib().mbt_ring_autotune(self.ptr, ctypes.byref(best), dptr(ms)))
        return best.value, {k: v for k, v in zip((1, 2, 4, 8, 16, 32, 64), ms.tolist()) if v > 0}

    def graph_builds(self):
        """Turn graphs captured and instantiated so far."""
        n = ctypes.c_long()
        _check(lib().mbt_ring_graph_builds(self.ptr, ctypes.byref(n)))
        return n.value

    def placement(self):
        """(coordinate allocations tried at set-up, stream GB/s of the one kept)"""
        n, g = ctypes.c_int(), ctypes.c_double()
        _check(lib().mbt_ring_placement(self.ptr, ctypes.byref(n), ctypes.byref(g)))
        return n.value, g.value

    def stream_ceiling(self, reps=10):
        """GB/s of a plain in-place update of the ring's own coordinate arrays
        (bit-preserving): the ceiling their placement sets for the turn pass."""
        out = ctypes.c_double()
        _check(lib().mbt_ring_stream_ceiling(self.ptr, reps, ctypes.byref(out)))
        return out.value


class Ensemble(Ring):
    """Independent rings tracked together on one GPU (mbt_ensemble_* C ABI).

    Bunch ordinals, stats and ampinv rows run over the rings' bunches in ring
    order; rf rows hold one (phai0, Vrf0) pair per ring."""

    def __init__(self, cfgs, device=0, init_dist=True, graphs=True, rw_gemm=True, mixed=False, rw_overlap=False,
                 host_init=False):
        L = lib()
        self.cfgs = list(cfgs)
        self.R = len(self.cfgs)
        self.nb = [c.geti("Nbunch") for c in self.cfgs]
        self.nps = [c.geti("Np") for c in self.cfgs]
        self.Nb = sum(self.nb)
        self.Nc = self.cfgs[0].geti("Ncell") if self.cfgs else 0
        self.bunch_np = np.repeat(self.nps, self.nb)
        self.first = np.concatenate([[0], np.cumsum(self.nb)])
        arr = (ctypes.c_void_p * self.R)(*[c.ptr.value for c in self.cfgs])
        p = ctypes.c_void_p()
        _check(L.mbt_ensemble_create(arr, self.R, device, _flags(init_dist, graphs, rw_gemm, mixed, rw_overlap,
                                                                 host_init=host_init),
                                     ctypes.byref(p)))
        self.ptr = p

    def size(self):
        r, g, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        _check(lib().mbt_ensemble_size(self.ptr, ctypes.byref(r), ctypes.byref(g), ctypes.byref(n)))
        return r.value, g.value, n.value

    def ring_rows(self, k, arr):
        """Rows of ring k in a per-bunch array [..., G, ...] (bunch axis -2)."""
        return arr[..., self.first[k]:self.first[k + 1], :]

    def set_bunch(self, g, soa6):
        a = np.ascontiguousarray(soa6, dtype=np.float64).reshape(-1)
        if a.size != 6 * self.bunch_np[g]:
            raise MbtError("bunch %d holds %d macroparticles" % (g, self.bunch_np[g]))
        _check(lib().mbt_ensemble_set_bunch(self.ptr, g, dptr(a)))

    def get_bunch(self, g):
        a = np.zeros(6 * int(self.bunch_np[g]))
        _check(lib().mbt_ensemble_get_bunch(self.ptr, g, dptr(a)))
        return a.reshape(6, -1)

    def track(self, nturns, want_stats=True, want_amp=True, want_rf=True):
        st = np.zeros((nturns, self.Nb, NSTAT)) if want_stats else None
        am = np.zeros((nturns, self.Nb, NAMP)) if want_amp else None
        rf = np.zeros((nturns, self.R, 2)) if want_rf else None
        _check(lib().mbt_ring_track(self.ptr, nturns, dptr(st), dptr(am), dptr(rf)))
        return st, am, rf


def ws_generate(cfg):
    """Host set-up of the weak-strong model (no device): slice centroids of every
    bucket [h][6][S], npop[S], grid counts [S][5][5], macro-electrons [6][Np]."""
    h, S, Np = cfg.geti("Nharm"), 2 * cfg.geti("Nslc") + 1, cfg.geti("Np")
    sl = np.zeros(h * 6 * S)
    npop = (ctypes.c_int * S)()
    nofe = (ctypes.c_int * (25 * S))()
    dist = np.zeros(6 * Np)
    _check(lib().mbt_ws_generate(cfg.ptr, dptr(sl), npop, nofe, dptr(dist)))
    return (sl.reshape(h, 6, S), np.array(list(npop)), np.array(list(nofe)).reshape(S, 5, 5),
            dist.reshape(6, Np))


class WeakStrong:
    """Weak-strong bunches with the fast beam-ion interaction on one GPU
    (mbt_ws_* C ABI)."""

    def __init__(self, cfg, device=0, bunch_currents=True, serial_train=False, mixed=False):
        L = lib()
        self.cfg = cfg
        self.Nb = cfg.geti("Nbunch")
        self.h = cfg.geti("Nharm")
        self.S = 2 * cfg.geti("Nslc") + 1
        p = ctypes.c_void_p()
        flags = ((MBT_WS_BUNCH_CURRENTS if bunch_currents else 0) | (MBT_WS_SERIAL_TRAIN if serial_train else 0)
                 | (MBT_WS_MIXED if mixed else 0))
        _check(L.mbt_ws_create(cfg.ptr, device, flags, ctypes.byref(p)))
        self.ptr = p

    def close(self):
        if getattr(self, "ptr", None):
            lib().mbt_ws_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    def track(self, nturns):
        cm = np.zeros((nturns, self.Nb, 4))
        mon = np.zeros((nturns, self.Nb, 18))
        ion = np.zeros((nturns, self.Nb, 3))
        zcm = np.zeros((nturns, self.h))
        _check(lib().mbt_ws_track(self.ptr, nturns, dptr(cm), dptr(mon), dptr(ion), dptr(zcm)))
        return cm, mon, ion, zcm

    def async_max(self):
        return lib().mbt_ws_async_max(self.ptr)

    def track_async(self, nturns):
        """Enqueues nturns turns (at most async_max()) without waiting."""
        _check(lib().mbt_ws_track_async(self.ptr, nturns))
        self._async_n = nturns

    def track_wait(self):
        """Waits for the turns of track_async; returns their outputs as track does."""
        n = getattr(self, "_async_n", 0)
        cm = np.zeros((n, self.Nb, 4))
        mon = np.zeros((n, self.Nb, 18))
        ion = np.zeros((n, self.Nb, 3))
        zcm = np.zeros((n, self.h))
        _check(lib().mbt_ws_track_wait(self.ptr, dptr(cm), dptr(mon), dptr(ion), dptr(zcm)))
        self._async_n = 0
        return cm, mon, ion, zcm

    def set_timing(self, on=True):
        _check(lib().mbt_ws_set_timing(self.ptr, 1 if on else 0))

    def train_time(self):
        """(milliseconds summed over the ion trains of the last track call, trains)"""
        ms, n = ctypes.c_double(), ctypes.c_long()
        _check(lib().mbt_ws_train_time(self.ptr, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def slices(self, ib):
        out = np.zeros(6 * self.S)
        _check(lib().mbt_ws_get_slices(self.ptr, ib, dptr(out)))
        return out.reshape(6, self.S)
