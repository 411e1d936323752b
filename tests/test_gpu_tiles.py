"""Bunches at the edges of the coordinate tiles (the device keeps every bunch's
six coordinates in tiles of 4,096 slots, each bunch starting on a tile): sizes
one below, at, one above and at twice a tile.  The device-generated
distributions read back bit for bit equal to the oracle's host-generated ones;
a modified bunch written and read back is the same bits; three turns of the
"wakes" configuration (broadband resonator, resistive wall, LR resonators) on
those bunches match the oracle to RTOL, statistics and every coordinate.

Coordinates are read between turns: the device applies a turn's
resistive-wall kick at the start of the next pass, the oracle within the turn,
so the read-back must add the kick still pending.  Writing a bunch back
unchanged between turns must drop that pending kick (else it is applied twice):
two more turns after such a write still match the oracle."""
import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from test_gpu_parity import RTOL, stat_errors
from variants import write_variant

pytestmark = pytest.mark.gpu


def coordinate_error(g, orc):
    """Largest |gpu - oracle| over every bunch, per coordinate scaled by that
    coordinate's largest magnitude in the bunch."""
    worst = 0.0
    for ib in range(g.Nb):
        cg_, co = g.get_bunch(ib), orc.get_bunch(ib)
        scale = np.maximum(np.max(np.abs(co), axis=1, keepdims=True), 1e-300)
        worst = max(worst, float(np.max(np.abs(cg_ - co) / scale)))
    return worst


@pytest.mark.parametrize("np_", [4095, 4096, 4097, 8192])
def test_tile_edge_bunches(tmp_path, np_):
    path = write_variant("wakes", str(tmp_path), distribution={"Np": np_})
    cfg = m.Config(path, str(tmp_path) + "/")
    print("case tile edge: Np", np_, flush=True)
    g = m.Ring(cfg)
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    for ib in range(g.Nb):
        assert np.array_equal(g.get_bunch(ib), orc.get_bunch(ib)), ib
    b = orc.get_bunch(g.Nb - 1).copy()
    b[2] *= 1.5  # x of the last bunch, through the copy-in
    g.set_bunch(g.Nb - 1, b)
    orc.set_bunch(g.Nb - 1, b)
    assert np.array_equal(g.get_bunch(g.Nb - 1), b)
    with pytest.raises(m.MbtError):
        g.set_bunch(0, b[:, :-1])  # one macroparticle short
    e0 = stat_errors(g.prepare(), orc.prepare()).max()
    gst, _, _ = g.track(3)
    ost, _, _ = orc.track(3)
    e = stat_errors(gst, ost).max()
    worst = coordinate_error(g, orc)
    # the same coordinates written back between turns change nothing
    for ib in (0, g.Nb - 1):
        g.set_bunch(ib, g.get_bunch(ib))
    gst2, _, _ = g.track(2)
    ost2, _, _ = orc.track(2)
    e2 = stat_errors(gst2, ost2).max()
    worst2 = coordinate_error(g, orc)
    print("Np %d: stats %.2g then %.2g (prepare %.2g), coordinates %.2g then %.2g"
          % (np_, e, e2, e0, worst, worst2), flush=True)
    assert e0 <= RTOL and e <= RTOL and e2 <= RTOL
    assert worst <= RTOL and worst2 <= RTOL
    g.close()
    orc.close()


def test_ensemble_of_tile_edge_rings(tmp_path):
    """Three rings of 4,095, 4,097 and 8,193 macroparticles per bunch in one
    ensemble: each ring's bunches start on tiles after the padded bunches of the
    rings before it.  Every ring matches the oracle to RTOL, statistics and the
    coordinates of its first and last bunch read between turns."""
    cfgs = []
    for k, np_ in enumerate((4095, 4097, 8193)):
        d = str(tmp_path / ("r%d" % k))
        path = write_variant("wakes", d, distribution={"Np": np_, "iseed_LON": 3 + k, "iseed_HOR": 23 + k,
                                                        "iseed_VER": 43 + k})
        cfgs.append(m.Config(path, d + "/"))
    print("case tile-edge ensemble", flush=True)
    ens = m.Ensemble(cfgs)
    ens.prepare()
    est, _, erf = ens.track(3)
    for k, cfg in enumerate(cfgs):
        orc = o.OracleRing(cfg, o.RNG_PHILOX)
        orc.prepare()
        ost, _, orf = orc.track(3)
        e = stat_errors(est[:, ens.first[k]:ens.first[k + 1]], ost).max()
        worst = 0.0
        for ib in (0, orc.Nb - 1):
            cg_, co = ens.get_bunch(ens.first[k] + ib), orc.get_bunch(ib)
            scale = np.maximum(np.max(np.abs(co), axis=1, keepdims=True), 1e-300)
            worst = max(worst, float(np.max(np.abs(cg_ - co) / scale)))
        print("ring %d (Np %d): stats %.2g, coordinates %.2g" % (k, orc.Np, e, worst), flush=True)
        assert e <= RTOL and worst <= RTOL
        assert (np.abs(erf[:, k] - orf) / np.abs(orf)).max() <= RTOL
        orc.close()
    ens.close()


@pytest.mark.parametrize("np_", [4097, 8193])
def test_quantum_excitation_across_tiles(tmp_path, np_):
    """Quantum excitation on bunches longer than a tile: each macroparticle's
    normals come from a counter-based stream keyed by its index in the bunch
    (not by its slot in the tiles), so they equal the oracle's bit for bit and
    the run meets RTOL, statistics and coordinates, on every tile of the
    bunch."""
    path = write_variant("qe", str(tmp_path), distribution={"Np": np_})
    cfg = m.Config(path, str(tmp_path) + "/")
    print("case qe across tiles: Np", np_, flush=True)
    g = m.Ring(cfg)
    orc = o.OracleRing(cfg, o.RNG_PHILOX)
    e0 = stat_errors(g.prepare(), orc.prepare()).max()
    gst, _, _ = g.track(4)
    ost, _, _ = orc.track(4)
    e = stat_errors(gst, ost).max()
    worst = 0.0
    for ib in range(g.Nb):
        cg_, co = g.get_bunch(ib), orc.get_bunch(ib)
        sc = np.maximum(np.std(co, axis=1, keepdims=True), 1e-300)
        worst = max(worst, float(np.max(np.abs(cg_ - co) / sc)))
    print("qe Np %d: stats %.2g (prepare %.2g), coordinates %.2g of the rms" % (np_, e, e0, worst), flush=True)
    assert e0 <= RTOL and e <= RTOL and worst <= RTOL
    g.close()
    orc.close()
