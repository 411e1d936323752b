"""The weak-strong oracle (oracle/oracle_ws.c) against the reference's own
functions, and the reference's Faddeeva function against the oracle's.

Fixtures (tests/golden/ws_*.npz, wofz_ref.npz) come from oracle/_ref/refwsharness
and oracle/_ref/refwofz, which link the reference's translation units built in
place (tests/golden/make_golden.py).  The bar is bitwise: per-turn bunch
centroids, monitor-turn statistics, ion-cloud sums after every bunch of the
train, bucket z-centroids, final slice arrays and the generated macro-electrons.
"""
import os

import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from variants import WS_VARIANTS, write_ws_variant

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_oracle(name, tmp_path):
    path, mode = write_ws_variant(name, str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    w = o.OracleWS(cfg, mode)
    n = cfg.geti("NrevTot")
    cm, mon, ion, zcm, rc = w.track(n)
    return cfg, w, cm, mon, ion, zcm, rc


@pytest.mark.parametrize("name", sorted(WS_VARIANTS))
def test_ws_oracle_bitwise_vs_reference(tmp_path, name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    cfg, w, cm, mon, ion, zcm, rc = run_oracle(name, tmp_path)
    assert rc == 0
    n, nb = cm.shape[0], cm.shape[1]
    assert np.array_equal(cm, g["cm"][:, 2:].reshape(n, nb, 4))
    assert np.array_equal(ion, g["ion"][:, 2:].reshape(n, nb, 3))
    monturns = [t for t in range(n) if t % cfg.geti("NrevMon") == 0]
    assert np.array_equal(mon[monturns].reshape(-1, 18), g["mon"][:, 2:])
    assert np.array_equal(zcm, g["zcm"][:, 1:])
    sl = np.concatenate([w.slices(ib).T for ib in range(nb)])
    assert np.array_equal(sl, g["slices"][:, 2:8])
    npop, nofe = w.population()
    assert np.array_equal(npop, g["slices"][: w.S, 8].astype(int))
    assert nofe.sum() == npop.sum()
    # generated macro-electrons, as the reference's manager writes them (%15.12g)
    d = g["dist"]
    mine = np.stack([w.dist[0, :64] / 1e-9, w.dist[1, :64], w.dist[2, :64] / 1e-3, w.dist[3, :64] / 1e-3,
                     w.dist[4, :64] / 1e-3, w.dist[5, :64] / 1e-3], axis=1)
    assert np.allclose(mine, d, rtol=1e-11, atol=0)
    w.close()


def test_ws_reference_currents_leave_ions_uncharged(tmp_path):
    """With the reference's STRONG-mode bunch currents (all zero) macro-ions
    carry no charge: every ion-cloud charge sum is zero and the FBII leaves the
    centroids exactly as without it."""
    _, w, cm, mon, ion, zcm, _ = run_oracle("ws_fbii_refcur", tmp_path)
    assert np.all(ion[..., 2] == 0.0)
    assert np.all(np.isfinite(cm))


def test_ws_ion_cloud_grows_by_one_bunch_of_ions(tmp_path):
    """After the k-th bunch of the train the cloud holds (k+1) bunches' ions:
    its total charge is (k+1) times one bunch's (equal currents)."""
    cfg, w, cm, mon, ion, zcm, _ = run_oracle("ws_fbii", tmp_path)
    q = ion[0, :, 2]
    assert q[0] > 0
    assert np.allclose(q, q[0] * np.arange(1, len(q) + 1), rtol=1e-12)


def test_ws_optics_conserves_invariant(tmp_path):
    """FBII and RW off: every slice centroid rotates on its Courant-Snyder
    ellipse, so gamma x^2 + 2 alpha x x' + beta x'^2 is conserved to round-off."""
    cfg, w, cm, mon, ion, zcm, _ = run_oracle("ws_optics", tmp_path)
    mon_t = [t for t in range(cm.shape[0]) if t % cfg.geti("NrevMon") == 0]
    amp = mon[mon_t][:, :, 12:15]
    assert np.allclose(amp[1:, :, 1:], amp[:1, :, 1:], rtol=1e-10)


def test_wofz_oracle_bitwise_vs_reference():
    g = np.load(os.path.join(GOLD, "wofz_ref.npz"))
    for x, y, u, v, f in zip(g["x"], g["y"], g["u"], g["v"], g["flag"]):
        uu, vv, ff = o.wofz(x, y)
        if f < 0:
            assert ff == 1
            continue
        assert ff == 0 and uu == u and vv == v, (x, y)


@pytest.mark.skipif(not os.path.exists(os.path.join(o.REF_DIR, "refwsharness")),
                    reason="reference harness not built (no /root/reference)")
def test_ws_live_harness_matches_fixture(tmp_path):
    """Where the reference is present, its harness reproduces the committed fixture."""
    import subprocess
    path, mode = write_ws_variant("ws_fbii", str(tmp_path))
    r = subprocess.run([os.path.join(o.REF_DIR, "refwsharness"), path, str(tmp_path), "0", str(mode)], capture_output=True)
    assert r.returncode == 0
    g = np.load(os.path.join(GOLD, "ws_fbii.npz"))
    assert np.array_equal(np.loadtxt(os.path.join(tmp_path, "cm.txt"), ndmin=2), g["cm"])


@pytest.mark.parametrize("name", ["ws_fbii", "ws_h416"])
def test_product_setup_equals_oracle(tmp_path, name):
    """The product's host set-up (mbt_ws_generate, an independent restatement)
    gives the oracle's slice centroids, populations, grid counts and
    macro-electrons bit for bit (no GPU needed)."""
    path, mode = write_ws_variant(name, str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    sl, npop, nofe, dist = m.ws_generate(cfg)
    w = o.OracleWS(cfg, mode)
    onpop, onofe = w.population()
    assert np.array_equal(npop, onpop) and np.array_equal(nofe, onofe)
    assert np.array_equal(dist, w.dist)
    for ib, kb in enumerate(cfg.buckets):
        assert np.array_equal(sl[kb], w.slices(ib)), (ib, kb)
    w.close()
