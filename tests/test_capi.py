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
