"""The C-ABI library loads without a GPU and exports every function that
include/mbtrack_mi355x.h declares."""
import os
import re

import mtrack_amd as m

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mbtrack_mi355x.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?\**\s*\b(mbt_[a-z0-9_]+)\s*\(", text, re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol():
    L = m.lib()
    names = declared()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_reporting():
    try:
        m.Config("/nonexistent/file.conf")
    except m.MbtError as e:
        assert "cannot open" in str(e)
    else:
        raise AssertionError("expected failure")


def test_ensemble_argument_checks(tmp_path):
    """Rejected before any device call: rings of one ensemble share the slicing."""
    from variants import write_variant

    a = m.Config(write_variant("wakes", str(tmp_path / "a"), selffield={"Ncell": 40}), str(tmp_path / "a") + "/")
    b = m.Config(write_variant("rsis", str(tmp_path / "b"), selffield={"Ncell": 60}), str(tmp_path / "b") + "/")
    try:
        m.Ensemble([a, b])
    except m.MbtError as e:
        assert "share Ncell" in str(e)
    else:
        raise AssertionError("expected failure")
    try:
        m.Ensemble([])
    except m.MbtError as e:
        assert "no ring" in str(e)
    else:
        raise AssertionError("expected failure")


def test_ring_rejects_invalid_configurations(tmp_path):
    """Configurations the turn cannot run are rejected with a message before any
    device call: no macroparticles, no monitor interval, wakes without slices,
    RF feedback without a longitudinal long-range resonator, a STRONG model."""
    from variants import write_variant

    cases = [("np0", "wakes", dict(distribution={"Np": 0}), "empty ring"),
             ("mon0", "wakes", dict(tracking={"NrevMon": 0}), "NrevMon"),
             ("noncell", "wakes", dict(selffield={"Ncell": 0}), "Ncell"),
             ("strong", "wakes", dict(tracking={"BunchModel": "STRONG"}), "weak-strong")]
    for name, variant, over, msg in cases:
        d = str(tmp_path / name)
        cfg = m.Config(write_variant(variant, d, **over), d + "/")
        try:
            m.Ring(cfg)
        except m.MbtError as e:
            assert msg in str(e), (name, str(e))
        else:
            raise AssertionError("%s: expected a rejection" % name)
