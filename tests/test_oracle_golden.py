"""The oracle (oracle/oracle.c) against the reference's own outputs.

Golden fixtures tests/golden/<variant>.npz were written by the reference's code
(oracle/_ref/refharness, see tests/golden/make_golden.py).  With the reference's
random stream (glibc rand + float polar Box-Muller, per-rank state) the oracle
must reproduce every statistic, RF value, ampinv and final coordinate BIT FOR
BIT, and the project's .conf reader must reproduce every derived parameter bit
for bit.  Where the reference harness is built (this container) the fixtures
are also regenerated live and compared, so stale fixtures fail.
"""
import os
import subprocess

import numpy as np
import pytest

import mtrack_amd as m
import oracle as o
from variants import VARIANTS, conf_template, write_variant

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def oracle_run(path, work, rng=o.RNG_REFERENCE):
    cfg = m.Config(path, work)
    r = o.OracleRing(cfg, rng)
    st0 = r.prepare()
    n = cfg.geti("NrevTot")
    st, am, rf = r.track(n)
    fin = r.final_stats() if cfg.geti("scan") == 0 else None
    return cfg, r, st0, st, am, rf, fin


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_oracle_bitwise_vs_reference(tmp_path, name):
    g = fixture(name)
    path = write_variant(name, str(tmp_path))
    text = conf_template(open(path).read(), str(tmp_path))
    assert text == str(g["conf"]), "variants.py changed: regenerate tests/golden"
    cfg, r, st0, st, am, rf, fin = oracle_run(path, str(tmp_path) + "/")
    nb = cfg.geti("Nbunch")
    gst = g["stats"]
    rows = [st0] + [st[t] for t in range(len(st))] + ([fin] if fin is not None else [])
    mine = np.concatenate(rows, axis=0)
    assert mine.shape == gst[:, 2:].shape
    assert np.array_equal(mine, gst[:, 2:]), "statistics differ from the reference"
    assert np.array_equal(rf, g["rf"][:, 1:])
    mon = (np.arange(len(st)) + 1) % cfg.geti("NrevMon") == 0
    assert np.array_equal(am[mon].reshape(-1, 6), g["ampinv"][:, 2:])
    parts = r.get_bunch(0)
    assert np.array_equal(parts.T, g["parts"])
    if cfg.geti("Ncell") > 0 and g["hist"].size:
        h = g["hist"]
        last = h[h[:, 0] == h[:, 0].max()]
        counts, trash = r.histograms()
        assert np.array_equal(trash, last[:, 2:4].astype(int)), "trash counters differ"
        nc = cfg.geti("Ncell")
        # interior bins: clamped rule == drop rule
        assert np.array_equal(counts[:, 1:nc - 1], last[:, 5:4 + nc - 1])
        assert np.all(counts.sum(1) == cfg.geti("Np"))
    assert nb == cfg.geti("Nbunch")


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_config_params_bitwise(tmp_path, name):
    g = fixture(name)
    path = write_variant(name, str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    dump = os.path.join(str(tmp_path), "mine.txt")
    cfg.dump(dump)
    mine = {}
    with open(dump) as f:
        for line in f:
            p = line.split()
            if len(p) == 2:
                try:
                    mine[p[0]] = float(p[1])
                except ValueError:
                    pass
    checked = 0
    for k, v in zip(g["param_names"], g["param_values"]):
        k = str(k)
        if k == "bunch_out":
            continue
        if k in mine:
            assert mine[k] == v or (np.isnan(v) and np.isnan(mine[k])), (k, mine[k], v)
            checked += 1
    assert checked > 60
    outs = [int(v) for k, v in zip(g["param_names"], g["param_values"]) if str(k) == "bunch_out"]
    assert outs == [int(line.split()[1]) for line in open(dump) if line.startswith("bunch_out")]


def test_product_generator_equals_oracle_generator(tmp_path):
    """Two independent restatements of weak_generate_bunch_distribution (the
    product's table walk and the oracle's closed-form window) agree bit for bit;
    the golden statistics pin both to the reference."""
    path = write_variant("fromfile", str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    for kb in cfg.buckets[:5]:
        assert np.array_equal(cfg.generate_bunch(kb), o.generate_bunch(cfg, kb))


@pytest.mark.skipif(not os.path.exists(o.REFHARNESS), reason="reference harness not built here")
@pytest.mark.parametrize("name", ["wakes", "lr_fb"])
def test_fixtures_are_current(tmp_path, name):
    g = fixture(name)
    path = write_variant(name, str(tmp_path))
    r = subprocess.run([o.REFHARNESS, path, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    st = np.loadtxt(os.path.join(str(tmp_path), "stats.txt"), ndmin=2)
    assert np.array_equal(st, g["stats"])


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_greens_product_equals_oracle(tmp_path, name):
    """The product's Green functions (mbt_greens_functions: resonators, short-range
    resistive wall with its tabulated first bins) equal the oracle's bit for bit;
    the oracle's are pinned to the reference through the bit-identical turns
    above."""
    path = write_variant(name, str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    if cfg.geti("Ncell") <= 0:
        pytest.skip("no selffield")
    gl, gv, gh, planes = cfg.greens()
    orc = o.OracleRing(cfg)
    ol, ov, oh, oplanes = orc.greens()
    assert np.array_equal(gl, ol) and np.array_equal(gv, ov) and np.array_equal(gh, oh), name
    assert planes == oplanes


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_oracle_greens_bitwise_vs_reference(tmp_path, name):
    """The oracle's Green functions (resonators, wake files, short-range
    resistive wall) equal the reference's own (selffield_model_t Gl/Gv/Gh after
    construct_greensfunc_resonator / _file / _RW, dumped by the reference
    harness) bit for bit, and so do the plane counters."""
    g = fixture(name)
    path = write_variant(name, str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    if cfg.geti("Ncell") <= 0:
        pytest.skip("no selffield")
    ol, ov, oh, planes = o.OracleRing(cfg).greens()
    ref = g["greens"]
    assert ref.shape == (cfg.geti("Ncell"), 4)
    assert np.all(np.isfinite(ref[:, 1:]))
    assert np.array_equal(ol, ref[:, 1]) and np.array_equal(ov, ref[:, 2]) and np.array_equal(oh, ref[:, 3]), name
    assert planes == list(g["greens_planes"])
