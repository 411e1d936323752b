"""Mixed precision (MBT_RING_MIXED, MBT_WS_MIXED) against FP64.

Mixed: the wake (selffield) convolutions in FP32 arithmetic with hi/lo
compensation, the resistive-wall sum as an FP32 MFMA GEMM with hi/lo operand
splitting and FP64 sums of 64-step FP32 chunks; for weak-strong bunches the
Faddeeva function of the field tables in FP32.  Coordinates, phase advance and
the LR / RF-feedback recurrence stay FP64, and both runs share the
counter-based QE stream, so every difference is the precision's.  Bounds: on
the small 64-bunch ring (stable, 2,000 turns) centroids within 1e-6 of the rms
and rms within 1e-6 relative; on the same ring driven unstable by the resistive
wall, the growth rate within 1 % of the FP64 run's.  (In the unstable run the
centroids themselves are not compared: the instability grows from noise-level
centroids, so any perturbation at that level reappears amplified.)  The FP64
GPU run is itself within 1e-12 of the oracle (tests/test_gpu_parity.py); the
first turns of the mixed run are also checked against the oracle directly.
Measured with the FP32 GEMM on the lone ring: centroids 1.0e-9 of the rms, rms
5.4e-11, growth rates 1.0e-8 relative.
"""
import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from mtrack_amd import confgen as cg
from test_gpu_parity import stat_errors
from variants import write_variant, write_ws_variant

pytestmark = pytest.mark.gpu
CENTROID_TOL = 1e-6
GROWTH_TOL = 0.01


def growth_rates(am, nrev_mon, first):
    """Per transverse plane: slope per turn of log(bunch-averaged centroid action)
    over the monitor turns from turn `first` on."""
    turns = np.arange(len(am))
    mon = (turns + 1) % nrev_mon == 0
    a = am[mon][..., 3:6].mean(1)
    t = turns[mon]
    sel = t >= first
    return np.array([np.polyfit(t[sel], np.log(a[sel, p]), 1)[0] for p in (1, 2)])


def _pair(cfg, turns):
    """FP64 and mixed runs of one ring; the mixed one with the resistive-wall sum
    as the FP32 hi/lo MFMA GEMM (MBT_RING_RW_GEMM: a lone ring otherwise keeps
    its per-bunch FP64 sum)."""
    runs = {}
    for mixed in (False, True):
        g = m.Ring(cfg, mixed=mixed, force_gemm=mixed)
        g.prepare()
        runs[mixed] = g.track(turns)
        g.close()
    return runs


def test_mixed_ring_centroids(tmp_path):
    """The small 64-bunch ring (64 x 1,000 macroparticles, QE on, copper RW,
    broadband resonator, 208 mA): 2,000 turns of FP64 and of mixed precision."""
    d = str(tmp_path)
    cfg = m.Config(cg.wl_c1(d, turns=2000), d + "/")
    print("case mixed ring c1", flush=True)
    runs = _pair(cfg, 2000)
    e = stat_errors(runs[True][0], runs[False][0])
    print("mixed vs fp64 over 2000 turns: centroid/rms error %.3g" % e.max(), flush=True)
    assert e.max() <= CENTROID_TOL


def test_mixed_ring_growth_rate(tmp_path):
    """The same ring driven unstable by the resistive wall (400 mA, 6 mm
    radius): 1,000 turns of FP64 and of mixed precision."""
    d = str(tmp_path)
    cfg = m.Config(cg.wl_c1(d, turns=1000, iring=400.0, beff=0.006), d + "/")
    print("case mixed ring c1 unstable", flush=True)
    runs = _pair(cfg, 1000)
    (s64, a64, _), (s32, a32, _) = runs[False], runs[True]
    e = stat_errors(s32, s64)
    rate64 = growth_rates(a64, 10, 500)
    rate32 = growth_rates(a32, 10, 500)
    rel = np.abs(rate32 - rate64) / np.abs(rate64)
    print("mixed vs fp64: centroid/rms error %.3g; growth rates fp64 %s mixed %s rel diff %s" % (
        e.max(), rate64, rate32, rel), flush=True)
    assert np.all(rate64 > 1e-3), "the configuration must be unstable"
    assert np.all(rel <= GROWTH_TOL)


def test_mixed_ring_first_turns_vs_oracle(tmp_path):
    d = str(tmp_path)
    cfg = m.Config(cg.wl_c1(d, turns=100), d + "/")
    print("case mixed ring c1 vs oracle", flush=True)
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    orc.prepare()
    ost, _, _ = orc.track(100)
    orc.close()
    g = m.Ring(cfg, mixed=True, force_gemm=True)
    g.prepare()
    gst, _, _ = g.track(100)
    g.close()
    e = stat_errors(gst, ost)
    print("mixed vs oracle over 100 turns: %.3g" % e.max(), flush=True)
    assert e.max() <= CENTROID_TOL


def test_mixed_ensemble_gemm(tmp_path):
    """Ten rings sharing one tap table: the FP32 hi/lo GEMM against the FP64
    per-bunch sum, and each ring's FP32 convolutions."""
    cfgs = []
    for k in range(10):
        dk = str(tmp_path / ("r%d" % k))
        path = write_variant("wakes", dk, ring={"Iring": 20.0 + 7.0 * k},
                             distribution={"Np": 400 + 50 * k, "iseed_LON": 11 + k, "iseed_HOR": 31 + k,
                                           "iseed_VER": 51 + k},
                             tracking={"NrevTot": 30})
        cfgs.append(m.Config(path, dk + "/"))
    print("case mixed ensemble", flush=True)
    a = m.Ensemble(cfgs, mixed=True)
    b = m.Ensemble(cfgs, rw_gemm=False)
    a.prepare()
    b.prepare()
    sa, _, _ = a.track(30)
    sb, _, _ = b.track(30)
    e = stat_errors(sa, sb)
    print("mixed ensemble vs fp64: %.3g" % e.max(), flush=True)
    assert e.max() <= CENTROID_TOL
    a.close()
    b.close()


def test_mixed_weakstrong(tmp_path):
    """FP32 Faddeeva function in the FBII field tables against FP64."""
    path, mode = write_ws_variant("ws_fbii", str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    print("case mixed weak-strong", flush=True)
    out = {}
    for mixed in (False, True):
        w = m.WeakStrong(cfg, bunch_currents=True, mixed=mixed)
        out[mixed] = w.track(30)
        w.close()
    errs = {}
    for k, lab in enumerate(("cm", "mon", "ion", "zcm")):
        r, g = out[False][k], out[True][k]
        errs[lab] = float(np.max(np.abs(g - r)) / max(np.max(np.abs(r)), 1e-300))
    print("mixed weak-strong vs fp64:", {k: "%.2g" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-5
