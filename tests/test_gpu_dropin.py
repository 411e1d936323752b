"""The command-line drop-in (mbtrack-mi355x, host C driver over the GPU ring)
against the unmodified reference program's output files for the same input
(tests/golden/e2e_wakes, written by mbtrack-mpi under MPICH with 21 ranks)."""
import os
import shutil
import subprocess

import pytest

import mtrack_amd as m
from filecmp_util import compare
from variants import write_variant

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "e2e_wakes")


def test_cli_files_match_reference(tmp_path):
    d = str(tmp_path)
    conf = write_variant("wakes", d)
    work = os.path.join(d, "work")
    os.makedirs(work)
    print("case e2e wakes", flush=True)
    r = subprocess.run([m.EXE, conf, work + "/"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ref_files = sorted(f for f in os.listdir(GOLD) if f != "stdout.txt")
    produced = set(os.listdir(work)) | {f for f in ("Ldist", "Hdist", "Vdist") if os.path.exists(os.path.join(d, f))}
    missing = [f for f in ref_files if f not in produced]
    assert not missing, missing
    bad = {}
    for f in ref_files:
        mine = os.path.join(work, f) if os.path.exists(os.path.join(work, f)) else os.path.join(d, f)
        errs = compare(open(os.path.join(GOLD, f)).read(), open(mine).read())
        if errs:
            bad[f] = errs[:3]
    assert not bad, bad
    out = r.stdout.replace("\r", "")
    import re
    out = re.sub(r"at time: [0-9:]+", "at time: T", out)
    assert compare(open(os.path.join(GOLD, "stdout.txt")).read(), out) == []


def test_cli_usage_error():
    r = subprocess.run([m.EXE], capture_output=True, text=True)
    assert r.returncode == 255 and "Usage" in r.stderr
