"""Ensembles of independent rings on one GPU (mbt_ensemble_create).

Every bunch of an ensemble runs through the same kernels, chunking and
reduction order as when its ring is tracked alone, so each ring of an ensemble
must reproduce its single-ring run bit for bit (statistics, ampinv, RF phase and
voltage, slice counts, coordinates); transitively it meets the single-ring
bound against the oracle, which is checked directly as well.  The members
differ in harmonic number, bunch count, macroparticle count, wakes, RF feedback,
current scan and quantum excitation.
"""
import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from test_gpu_parity import RTOL, stat_errors
from variants import write_variant

pytestmark = pytest.mark.gpu
MEMBERS = ["wakes", "lr_fb", "scan", "lr_hv", "fromfile"]
TURNS = 30
COMMON = dict(selffield={"Ncell": 40}, tracking={"NrevMon": 10})


def _configs(tmp_path):
    cfgs = []
    for name in MEMBERS:
        d = str(tmp_path / name)
        cfgs.append(m.Config(write_variant(name, d, **COMMON), d + "/"))
    return cfgs


def test_ensemble_equals_single_rings(tmp_path):
    cfgs = _configs(tmp_path)
    print("ensemble of", MEMBERS, flush=True)
    ens = m.Ensemble(cfgs)
    R, G, N = ens.size()
    assert R == len(MEMBERS) and G == sum(c.geti("Nbunch") for c in cfgs)
    assert N == sum(c.geti("Nbunch") * c.geti("Np") for c in cfgs)
    e0 = ens.prepare()
    est, eam, erf = ens.track(TURNS)
    ehc, etr = ens.histograms()
    for k, cfg in enumerate(cfgs):
        print("member", MEMBERS[k], flush=True)
        g = m.Ring(cfg)
        # the default chunk depends on the macroparticle count: use the ensemble's
        g.set_chunk(ens.chunk())
        g0 = g.prepare()
        gst, gam, grf = g.track(TURNS)
        ghc, gtr = g.histograms()
        sl = slice(ens.first[k], ens.first[k + 1])
        assert np.array_equal(e0[sl], g0), MEMBERS[k]
        assert np.array_equal(est[:, sl], gst), MEMBERS[k]
        assert np.array_equal(eam[:, sl], gam), MEMBERS[k]
        assert np.array_equal(erf[:, k], grf), MEMBERS[k]
        assert np.array_equal(ehc[sl], ghc) and np.array_equal(etr[sl], gtr), MEMBERS[k]
        for ib in (0, cfg.geti("Nbunch") - 1):
            assert np.array_equal(ens.get_bunch(ens.first[k] + ib), g.get_bunch(ib))
        g.close()
    ens.close()


def test_ensemble_matches_oracle(tmp_path):
    cfgs = _configs(tmp_path)
    ens = m.Ensemble(cfgs)
    ens.prepare()
    est, _, erf = ens.track(TURNS)
    for k, cfg in enumerate(cfgs):
        orc = o.OracleRing(cfg, o.RNG_PHILOX)
        orc.prepare()
        ost, _, orf = orc.track(TURNS)
        e = stat_errors(est[:, ens.first[k]:ens.first[k + 1]], ost)
        print("%s: turn err max %.3g" % (MEMBERS[k], e.max()), flush=True)
        assert e.max() <= RTOL
        assert (np.abs(erf[:, k] - orf) / np.abs(orf)).max() <= RTOL
        orc.close()
    ens.close()



def test_rw_gemm_equals_per_bunch_sum(tmp_path):
    """Ten rings sharing one resistive-wall tap table (same h, Nmlt, T0; their
    own currents, bunch counts and seeds): the ensemble's RW sum runs as one FP64
    MFMA GEMM.  It must agree with the per-bunch sum (MBT_RING_NO_RW_GEMM) to the
    rounding of the two summation orders, and each ring with the oracle."""
    cfgs = []
    for k in range(10):
        d = str(tmp_path / ("r%d" % k))
        path = write_variant("wakes", d, ring={"Iring": 20.0 + 7.0 * k},
                             distribution={"Np": 400 + 50 * k, "iseed_LON": 11 + k, "iseed_HOR": 31 + k,
                                           "iseed_VER": 51 + k},
                             tracking={"NrevTot": TURNS})
        cfgs.append(m.Config(path, d + "/"))
    a = m.Ensemble(cfgs)
    b = m.Ensemble(cfgs, rw_gemm=False)
    a.prepare()
    b.prepare()
    sa, _, _ = a.track(TURNS)
    sb, _, _ = b.track(TURNS)
    e = stat_errors(sa, sb)
    print("gemm vs per-bunch RW: max normalised difference %.3g" % e.max(), flush=True)
    assert e.max() <= RTOL
    for k, cfg in enumerate(cfgs[:3]):
        orc = o.OracleRing(cfg, o.RNG_PHILOX)
        orc.prepare()
        ost, _, _ = orc.track(TURNS)
        assert stat_errors(sa[:, a.first[k]:a.first[k + 1]], ost).max() <= RTOL
        # coordinates read between turns carry the GEMM path's pending kick
        for ib in (0, orc.Nb - 1):
            co = orc.get_bunch(ib)
            sc = np.maximum(np.std(co, axis=1, keepdims=True), 1e-300)
            assert (np.abs(a.get_bunch(a.first[k] + ib) - co) / sc).max() <= RTOL, (k, ib)
        orc.close()
    a.close()
    b.close()
