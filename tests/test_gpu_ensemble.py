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

