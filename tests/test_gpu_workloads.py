"""The benchmark workloads themselves against the CPU oracle, at the paths and
work decompositions the benchmark runs them with.

* Production ring (416 bunches in h = 416, 100 slices, broadband resonator,
  resistive wall over 50 past turns = 21,216 taps per bunch and plane, QE off):
  at 2,000 macroparticles per bunch for 30 turns with the split resistive-wall
  sum (earlier turns' rows in the leading blocks of the map, this turn's row
  after the reduction) and 8 items per lane, the decomposition the autotuner
  picks for the full ring; and the full ring (416 x 100,000) for 3 turns.
* Fast beam-ion train: 356 bunches + 60-bucket gap, 41 slices, 1,025 macro-ions
  per bunch passage, 2 turns, both forms of the ion train.
* Batched ensemble: the benchmark's 256 seeded rings (bunch counts 32-416,
  currents 0.1-5 mA, chromaticities -1..3, RF beam-loading feedback, resistive
  wall as one FP64 MFMA GEMM over all rings) with macroparticle counts drawn in
  100-1,000, 20 turns; 16 rings compared with the oracle, among them the
  smallest and the largest.

Bounds as in test_gpu_parity.py (centroids normalised by the rms, rms relative,
RTOL = 1e-12; slice counts and trash counters bit-exact); the beam-ion train
as test_gpu_ws.py measures it, against RTOL = 1e-12 (measured 5e-13, the cloud's
charge sum over 365,000 ions in a tree order against the reference's loop).
"""
import os

import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from mtrack_amd import confgen as cg
from test_gpu_parity import RTOL, stat_errors

pytestmark = pytest.mark.gpu


def _production(tmp_path, np_, turns):
    d = str(tmp_path)
    return m.Config(cg.write(os.path.join(d, "production.conf"), cg.wl_production(turns=turns, np_=np_)), d + "/")


def _compare_ring(g, orc, turns, label):
    gst, gam, grf = g.track(turns)
    ost, oam, orf = orc.track(turns)
    e = stat_errors(gst, ost)
    er = float(np.max(np.abs(grf - orf) / np.abs(orf)))
    print("%s: turn err max %.3g (per column %s), rf %.3g" % (
        label, e.max(), np.array2string(e.reshape(-1, 12).max(0), precision=2), er), flush=True)
    assert e.max() <= RTOL
    assert er <= RTOL
    hc_o, tr_o = orc.histograms()
    hc_g, tr_g = g.histograms()
    assert np.array_equal(hc_o, hc_g), "slice counts differ"
    assert np.array_equal(tr_o, tr_g), "trash counters differ"
    return gst, gam


def test_production_shape_split_rw(tmp_path):
    turns = 30
    cfg = _production(tmp_path, 2000, turns)
    print("case production shape: 416 x 2000, split RW sum, 8 items per lane, %d turns" % turns, flush=True)
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    g = m.Ring(cfg, rw_overlap=True)
    g.set_chunk(8)
    assert g.chunk() == 8
    e0 = stat_errors(g.prepare(), orc.prepare()).max()
    assert e0 <= RTOL
    _compare_ring(g, orc, turns, "production 416x2000")
    g.close()
    orc.close()


@pytest.mark.timeout(600)
def test_production_full_size(tmp_path):
    turns = 3
    cfg = _production(tmp_path, 100000, turns)
    print("case production full size: 416 x 100000, %d turns" % turns, flush=True)
    g = m.Ring(cfg)
    g0 = g.prepare()
    best, _ = g.autotune()
    print("autotuner chose %d items per lane" % best, flush=True)
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    e0 = stat_errors(g0, orc.prepare()).max()
    assert e0 <= RTOL
    _compare_ring(g, orc, turns, "production 416x100000")
    fo, fg = orc.final_stats(), g.final_stats()
    assert stat_errors(fg, fo).max() <= RTOL
    g.close()
    orc.close()


def _scaled_err(g, r):
    g, r = np.asarray(g), np.asarray(r)
    return float(np.max(np.abs(g - r)) / max(np.max(np.abs(r)), 1e-300))


@pytest.fixture(scope="module")
def fbii_oracle(tmp_path_factory):
    import bench
    d = str(tmp_path_factory.mktemp("fbii"))
    cfg = m.Config(bench.fbii_conf(d, turns=2), d + "/")
    print("oracle: 356-bunch beam-ion train, 2 turns", flush=True)
    orc = o.OracleWS(cfg, o.WS_IBUNCH_FILL)
    out = orc.track(2)
    assert out[-1] == 0
    sl = np.stack([orc.slices(ib) for ib in range(orc.Nb)])
    orc.close()
    return cfg, out[:4], sl


@pytest.mark.timeout(900)
@pytest.mark.parametrize("serial", [False, True], ids=["ion_parallel", "relay"])
def test_fbii_train_356(fbii_oracle, serial):
    from test_gpu_ws import RTOL as WS_RTOL
    cfg, (ocm, omon, oion, ozcm), sl_o = fbii_oracle
    print("case fbii 356-bunch train,", "relay" if serial else "ion-parallel", flush=True)
    g = m.WeakStrong(cfg, bunch_currents=True, serial_train=serial)
    gcm, gmon, gion, gzcm = g.track(2)
    sl_g = np.stack([g.slices(ib) for ib in range(g.Nb)])
    g.close()
    errs = {"cm": _scaled_err(gcm, ocm), "zcm": _scaled_err(gzcm, ozcm)}
    for k, lab in enumerate(["pos", "pos_sigma", "slope", "slope_sigma", "ampinv", "ampinv_cm"]):
        errs[lab] = _scaled_err(gmon[0][..., 3 * k:3 * k + 3], omon[0][..., 3 * k:3 * k + 3])
    for k, lab in enumerate(["ion_x", "ion_y", "ion_q"]):
        errs[lab] = _scaled_err(gion[..., k], oion[..., k])
    for k, lab in enumerate(["tau", "delta", "x", "xp", "z", "zp"]):
        errs["slice_" + lab] = _scaled_err(sl_g[:, k], sl_o[:, k]) if np.any(sl_o[:, k]) else \
            float(np.max(np.abs(sl_g[:, k])))
    print({k: "%.2g" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= WS_RTOL, errs


@pytest.mark.timeout(600)
def test_ensemble_256_rings(tmp_path):
    turns = 20
    d = str(tmp_path)
    paths = cg.wl_ensemble(d, nrings=256, np_range=(100, 1000), turns=turns)
    cfgs = [m.Config(p, d + "/") for p in paths]
    n = np.array([c.geti("Nbunch") * c.geti("Np") for c in cfgs])
    print("case ensemble: 256 rings, %d bunches, %d macroparticles, %d turns" % (
        sum(c.geti("Nbunch") for c in cfgs), n.sum(), turns), flush=True)
    ens = m.Ensemble(cfgs)
    ens.prepare()
    best, _ = ens.autotune()
    print("autotuner chose %d items per lane" % best, flush=True)
    est, _, erf = ens.track(turns)
    ehc, etr = ens.histograms()
    ens.close()
    order = np.argsort(n)
    pick = sorted(set([int(order[0]), int(order[-1])] + [int(order[i]) for i in np.linspace(1, 254, 14).astype(int)]))
    worst = 0.0
    for k in pick:
        orc = o.OracleRing(cfgs[k], o.RNG_PHILOX)
        orc.prepare()
        ost, _, orf = orc.track(turns)
        sl = slice(int(np.sum([c.geti("Nbunch") for c in cfgs[:k]])), None)
        sl = slice(sl.start, sl.start + cfgs[k].geti("Nbunch"))
        e = stat_errors(est[:, sl], ost).max()
        er = float(np.max(np.abs(erf[:, k] - orf) / np.abs(orf)))
        hc_o, tr_o = orc.histograms()
        orc.close()
        print("ring %d (%d x %d): turn err %.3g, rf %.3g" % (k, cfgs[k].geti("Nbunch"), cfgs[k].geti("Np"), e, er),
              flush=True)
        worst = max(worst, e, er)
        assert np.array_equal(ehc[sl], hc_o) and np.array_equal(etr[sl], tr_o), k
    assert worst <= RTOL
