"""GPU turn vs the CPU oracle on the same seeded inputs, through the C ABI.

Tolerances (FP64 everywhere on both sides):
  * slice counts and trash counters: bit-exact;
  * bunch rms and ampinv: relative error <= RTOL;
  * centroids: |gpu - oracle| <= RTOL * rms of the same coordinate (centroids of
    a centred bunch are ~1e-3 sigma, so a relative error of the centroid itself is
    not meaningful);
  * RF phase / voltage: relative error <= RTOL;
  * final coordinates of the first and last bunch: |gpu - oracle| <= RTOL * rms
    of the same coordinate.
The residual difference comes from the device's sin/cos (ocml, <= 1 ulp) against
glibc's, and from the summation order of moments and of the RW sum.  The QE
normals are bit-identical on both sides (counter-based stream, explicit-FMA
float Box-Muller), so QE-on runs are held to the same bar.
"""
import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from variants import VARIANTS, write_variant

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def stat_errors(g, r):
    """Per-column normalised error of stats arrays [..., 12].  Where the rms is
    zero (one macroparticle per bunch) the centroid is compared with the largest
    centroid of its column over the run."""
    g = np.asarray(g)
    r = np.asarray(r)
    err = np.zeros(g.shape)
    for p in range(3):
        for mean_c, sig_c in ((p, 3 + p), (6 + p, 9 + p)):
            scale = np.abs(r[..., sig_c])
            scale = np.where(scale > 0, scale, np.max(np.abs(r[..., mean_c])))
            scale = np.maximum(scale, 1e-300)
            err[..., mean_c] = np.abs(g[..., mean_c] - r[..., mean_c]) / scale
            err[..., sig_c] = np.abs(g[..., sig_c] - r[..., sig_c]) / scale
    return err


def run_pair(path, work, turns=None):
    cfg = m.Config(path, work)
    n = turns or cfg.geti("NrevTot")
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    ost0 = orc.prepare()
    ost, oam, orf = orc.track(n)
    g = m.Ring(cfg)
    gst0 = g.prepare()
    gst, gam, grf = g.track(n)
    return cfg, (orc, ost0, ost, oam, orf), (g, gst0, gst, gam, grf)


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_turns_match_oracle(tmp_path, name):
    path = write_variant(name, str(tmp_path))
    print("case", name, flush=True)
    cfg, (orc, ost0, ost, oam, orf), (g, gst0, gst, gam, grf) = run_pair(path, str(tmp_path) + "/")
    e0 = stat_errors(gst0, ost0).max()
    e = stat_errors(gst, ost)
    print("%s: init err %.3g, turn err max %.3g (per column %s)" % (
        name, e0, e.max(), np.array2string(e.reshape(-1, 12).max(0), precision=2)))
    assert e0 <= RTOL
    assert e.max() <= RTOL
    mon = (np.arange(len(ost)) + 1) % cfg.geti("NrevMon") == 0
    a_o, a_g = oam[mon], gam[mon]
    ea = np.abs(a_g - a_o) / np.maximum(np.abs(a_o), 1e-300)
    ea[np.abs(a_o) == 0] = np.abs(a_g[np.abs(a_o) == 0])
    print("%s: ampinv err %.3g" % (name, ea.max()))
    assert ea.max() <= RTOL
    er = np.abs(grf - orf) / np.abs(orf)
    assert er.max() <= RTOL
    # the end-of-run RF state is the one the last turn's map used
    ph, vrf, fb_on = g.rf_state()
    assert (ph, vrf) == (grf[-1, 0], grf[-1, 1])
    assert not fb_on or cfg.geti("has_rf_feedback")
    if cfg.geti("Ncell") > 0:
        hc_o, tr_o = orc.histograms()
        hc_g, tr_g = g.histograms()
        assert np.array_equal(hc_o, hc_g), "slice counts differ"
        assert np.array_equal(tr_o, tr_g), "trash counters differ"
    fo = orc.final_stats()
    fg = g.final_stats()
    assert stat_errors(fg, fo).max() <= RTOL
    # coordinates themselves
    ce = 0.0
    for ib in (0, g.Nb - 1):
        co = orc.get_bunch(ib)
        cg_ = g.get_bunch(ib)
        sc = np.std(co, axis=1, keepdims=True)
        sc = np.maximum(np.where(sc > 0, sc, np.abs(co).max(axis=1, keepdims=True)), 1e-300)
        ce = max(ce, float((np.abs(cg_ - co) / sc).max()))
    print("%s: coordinate err %.3g" % (name, ce))
    assert ce <= RTOL
    g.close()


def test_deterministic_bitwise(tmp_path):
    """Two GPU runs of the same input are bit-identical (exact integer histogram
    sums, fixed-order moment reductions)."""
    path = write_variant("wakes", str(tmp_path))
    out = []
    for _ in range(2):
        cfg = m.Config(path, str(tmp_path) + "/")
        g = m.Ring(cfg)
        g.prepare()
        st, am, rf = g.track(20)
        out.append((st, am, g.get_bunch(3)))
        g.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("items", [1, 2, 4, 8, 16, 32, 64])
def test_chunk_sizes_match_oracle(tmp_path, items):
    """Any work decomposition of the per-macroparticle pass gives the oracle's
    turns to the same bar (only the grouping of the moment sums changes); the
    autotuner leaves the state untouched and picks one of its candidates."""
    path = write_variant("wakes", str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    n = cfg.geti("NrevTot")
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    orc.prepare()
    ost, _, _ = orc.track(n)
    g = m.Ring(cfg)
    g.prepare()
    best, ms = g.autotune()
    assert best in ms and all(v > 0 for v in ms.values())
    assert set(ms) == {1, 2, 4, 8, 16, 32, 64}  # a small ring tries the 1- and 2-item chunks too
    g.set_chunk(items)
    assert g.chunk() == items
    gst, _, _ = g.track(n)
    assert stat_errors(gst, ost).max() <= RTOL
    hc_o, tr_o = orc.histograms()
    hc_g, tr_g = g.histograms()
    assert np.array_equal(hc_o, hc_g) and np.array_equal(tr_o, tr_g)


def test_qe_statistics_match_reference_stream(tmp_path):
    """Quantum excitation over 1,200 turns of relaxation (tests/test_qe_statistics.py
    explains the configuration and tolerances): the GPU run follows the oracle's
    counter-based stream turn by turn, and its sigma^2 curves meet the statistical
    bounds against the reference's own glibc stream."""
    from test_qe_statistics import CURVE_TOL, EQ_TOL
    from variants import QE_STAT_TURNS, qe_stat_compare, sigma2_curves, write_qe_stat
    d = str(tmp_path)
    cfg = m.Config(write_qe_stat(d), d + "/")
    print("case qe statistics", flush=True)
    g = m.Ring(cfg)
    g.prepare()
    gst, _, _ = g.track(QE_STAT_TURNS)
    g.close()
    sg = sigma2_curves(gst)
    curves = {}
    for name, rng in (("philox", o.RNG_PHILOX), ("reference", o.RNG_REFERENCE)):
        orc = o.OracleRing(cfg, rng)
        orc.prepare()
        st, _, _ = orc.track(QE_STAT_TURNS)
        orc.close()
        curves[name] = sigma2_curves(st)
    same = np.abs(sg - curves["philox"]) / curves["philox"]
    eq, curve = qe_stat_compare(curves["reference"], sg)
    print("gpu vs oracle (same stream) max rel %.3g; vs reference stream: equilibrium %s per-turn %s" % (
        same.max(), np.round(eq, 4), np.round(curve, 4)), flush=True)
    assert same.max() <= 1e-9
    assert np.all(eq <= EQ_TOL) and np.all(curve <= CURVE_TOL)


@pytest.mark.parametrize("name", ["wakes", "fromfile", "scan", "edge_single"])
def test_rw_overlap_matches_oracle(tmp_path, name):
    """The split resistive-wall sum (the taps of earlier turns' rows in the
    leading blocks of the turn's map, the fresh row's taps after the reduction;
    what rings of >= 4,194,304 macroparticles run) gives the oracle's turns, and
    replaying the turns as graphs equals launching them eagerly bit for bit."""
    path = write_variant(name, str(tmp_path))
    print("case rw overlap", name, flush=True)
    cfg = m.Config(path, str(tmp_path) + "/")
    n = cfg.geti("NrevTot")
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    orc.prepare()
    ost, oam, _ = orc.track(n)
    runs = []
    for graphs in (True, False):
        g = m.Ring(cfg, graphs=graphs, rw_overlap=True)
        g.prepare()
        gst, gam, grf = g.track(n)
        hc, tr = g.histograms()
        runs.append((gst, gam, grf, hc, tr, g.get_bunch(g.Nb - 1)))
        g.close()
    gst, gam = runs[0][0], runs[0][1]
    e = stat_errors(gst, ost)
    print("%s: rw overlap turn err max %.3g" % (name, e.max()), flush=True)
    assert e.max() <= RTOL
    mon = (np.arange(len(ost)) + 1) % cfg.geti("NrevMon") == 0
    ea = np.abs(gam[mon] - oam[mon]) / np.maximum(np.abs(oam[mon]), 1e-300)
    assert ea.max() <= RTOL
    hc_o, tr_o = orc.histograms()
    assert np.array_equal(runs[0][3], hc_o) and np.array_equal(runs[0][4], tr_o)
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def test_stream_ceiling_leaves_state(tmp_path):
    """The placement-ceiling probe rewrites every coordinate bit for bit: the
    ring's bunches and its further turns equal those of a ring never probed."""
    path = write_variant("wakes", str(tmp_path))
    print("case stream ceiling", flush=True)
    out = []
    for probe in (False, True):
        cfg = m.Config(path, str(tmp_path) + "/")
        g = m.Ring(cfg)
        g.prepare()
        g.track(5)
        if probe:
            gbps = g.stream_ceiling(3)
            print("ceiling %.1f GB/s" % gbps, flush=True)
            assert gbps > 0
        st, am, _ = g.track(10)
        out.append((st, am, g.get_bunch(0), g.get_bunch(g.Nb - 1)))
        g.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["wakes", "fromfile", "lr_hv", "edge_single", "ahc_amp"])
def test_device_generator_equals_host_generator(tmp_path, name):
    """Initial distributions made on the device (the default where the
    distribution is the reference's Gaussian generator) equal the host
    generator's bit for bit -- the host generator being pinned to the
    reference's weak_generate_bunch_distribution through the oracle's golden
    fixtures -- for every bunch and coordinate."""
    path = write_variant(name, str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    print("case device generator", name, flush=True)
    a = m.Ring(cfg)
    b = m.Ring(cfg, host_init=True)
    for ib in range(a.Nb):
        assert np.array_equal(a.get_bunch(ib), b.get_bunch(ib)), (name, ib)
    assert np.array_equal(a.get_bunch(0), cfg.generate_bunch(cfg.buckets[0]).reshape(6, -1))
    a.close()
    b.close()
