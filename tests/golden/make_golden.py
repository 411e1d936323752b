"""Writes the golden fixtures tests/golden/<variant>.npz from the reference's own
code: oracle/_ref/refharness (reference translation units compiled from
/root/reference/src by oracle/Makefile) run on every configuration of
tests/variants.py.  Run in the build container only:

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds the .conf text it was made from, the per-turn per-bunch
statistics (turn -1 = initial distribution, last row block = final statistics),
RF phase/voltage per turn, ampinv on monitor turns, the slice histogram of
selected turns, the final coordinates of the first bunch, and the derived
machine parameters.  All values are the reference's doubles printed %.17g.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from variants import VARIANTS, conf_template, write_e2e_variant, write_variant  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "refharness")


def load_params(path):
    names, vals = [], []
    with open(path) as f:
        for line in f:
            k, v = line.split()
            names.append(k)
            vals.append(float(v))
    return np.array(names), np.array(vals)


def make(name):
    with tempfile.TemporaryDirectory() as d:
        conf = write_variant(name, d)
        text = conf_template(open(conf).read(), d)
        first = None
        r = subprocess.run([HARNESS, conf, d, "0", "1", "-1"], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(r.stderr)
        st = np.loadtxt(os.path.join(d, "stats.txt"), ndmin=2)
        first = int(st[0, 1])
        r = subprocess.run([HARNESS, conf, d, "0", "1", str(first)], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(r.stderr)
        st = np.loadtxt(os.path.join(d, "stats.txt"), ndmin=2)
        rf = np.loadtxt(os.path.join(d, "rf.txt"), ndmin=2)
        am = np.loadtxt(os.path.join(d, "ampinv.txt"), ndmin=2)
        hp = os.path.join(d, "hist.txt")
        hist = np.loadtxt(hp, ndmin=2) if os.path.exists(hp) and os.path.getsize(hp) else np.zeros((0, 4))
        if hist.size:
            keep = (hist[:, 0] % 10 == 9) | (hist[:, 0] == hist[:, 0].max())
            hist = hist[keep]
        parts = np.loadtxt(os.path.join(d, "parts_%d.txt" % first), ndmin=2)
        pn, pv = load_params(os.path.join(d, "params.txt"))
        gp = os.path.join(d, "greens.txt")
        greens, planes = np.zeros((0, 4)), np.zeros(3, dtype=int)
        if os.path.exists(gp):
            lines = open(gp).read().split("\n")
            greens = np.array([[float(v) for v in l.split()] for l in lines if l and not l.startswith("planes")])
            planes = np.array([int(v) for v in [l for l in lines if l.startswith("planes")][0].split()[1:]])
    out = os.path.join(HERE, name + ".npz")
    np.savez_compressed(out, conf=np.array(text), stats=st, rf=rf, ampinv=am, hist=hist, parts=parts,
                        param_names=pn, param_values=pv, greens=greens, greens_planes=planes)
    return out




def make_inp():
    """Derived parameters of the reference's read_input on the legacy .inp form
    (tests/variants.py INP_LINES), from the harness with 0 turns."""
    from variants import write_inp
    with tempfile.TemporaryDirectory() as d:
        path = write_inp(d)
        r = subprocess.run([HARNESS, path, d, "0"], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(r.stderr)
        pn, pv = load_params(os.path.join(d, "params.txt"))
    out = os.path.join(HERE, "inp_params.npz")
    np.savez_compressed(out, param_names=pn, param_values=pv)
    return out


# per-variant section overrides of the end-to-end cases (none at present)
E2E_CASES = ["wakes", "lr_fb", "rsis", "ws_uniform", "rw_short", "wakefile", "wakes_blocks", "wakes_mon300"]

MPI_BIN = os.path.join(ROOT, "oracle", "_ref", "mbtrack-mpi")
MPIEXEC = "/opt/conda/bin/mpiexec"


def make_e2e(name="wakes"):
    """Output files of the unmodified reference program (mbtrack-mpi under
    MPICH, one rank per bunch + manager) for one variant, kept under
    tests/golden/e2e_<name>/ for the drop-in driver's file-level comparison."""
    import shutil
    out = os.path.join(HERE, "e2e_" + name)
    if name.startswith("ws_"):
        return make_e2e_ws(name)
    with tempfile.TemporaryDirectory() as d:
        conf = write_e2e_variant(name, d)
        work = os.path.join(d, "work")
        os.makedirs(work)
        st = np.loadtxt  # noqa: F841
        import re
        nb = int(re.search(r"Nbunch (\S+)", open(os.path.join(d, "params.txt")).read()).group(1)) \
            if os.path.exists(os.path.join(d, "params.txt")) else None
        if nb is None:
            r = subprocess.run([HARNESS, conf, d, "1"], capture_output=True, text=True)
            nb = int(float(re.search(r"^Nbunch (\S+)$", open(os.path.join(d, "params.txt")).read(), re.M).group(1)))
        cmd = "ulimit -s unlimited; %s -usize %d -n %d %s %s %s/" % (MPIEXEC, nb + 1, nb + 1, MPI_BIN, conf, work)
        r = subprocess.run(["bash", "-c", cmd], capture_output=True, text=True, cwd=d, timeout=600)
        if r.returncode:
            raise RuntimeError(r.stderr)
        if os.path.exists(out):
            shutil.rmtree(out)
        os.makedirs(out)
        for f in sorted(os.listdir(work)):
            if f.endswith(".log"):
                continue
            shutil.copy(os.path.join(work, f), os.path.join(out, f))
        for f in ("Ldist", "Hdist", "Vdist"):
            if os.path.exists(os.path.join(d, f)):
                shutil.copy(os.path.join(d, f), os.path.join(out, f))
        with open(os.path.join(out, "stdout.txt"), "w") as f:
            f.write(re.sub(r"at time: [0-9:]+", "at time: T", r.stdout))
    return out


WS_HARNESS = os.path.join(ROOT, "oracle", "_ref", "refwsharness")


def make_e2e_ws(name):
    """The weak-strong manager's files for a STRONG configuration, written by
    the reference's own writer functions through refwsharness (the MPI program
    cannot run this mode below h = 500, see oracle/ref_ws_harness.c)."""
    import re
    import shutil
    from variants import write_ws_variant
    out = os.path.join(HERE, "e2e_" + name)
    keep_out = {"cm.txt", "mon.txt", "ion.txt", "zcm.txt", "slices.txt"}
    with tempfile.TemporaryDirectory() as d:
        conf, mode = write_ws_variant(name, d)
        work = os.path.join(d, "work")
        os.makedirs(work)
        r = subprocess.run([WS_HARNESS, conf, work, "0", str(mode), "1"], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(r.stderr)
        if os.path.exists(out):
            shutil.rmtree(out)
        os.makedirs(out)
        for f in sorted(os.listdir(work)):
            if f not in keep_out:
                shutil.copy(os.path.join(work, f), os.path.join(out, f))
        with open(os.path.join(out, "stdout.txt"), "w") as f:
            f.write(re.sub(r"at time: [0-9:]+", "at time: T", r.stdout))
    return out
WOFZ_DUMP = os.path.join(ROOT, "oracle", "_ref", "refwofz")


def make_ws(name):
    """Weak-strong fixture from the reference's own functions
    (oracle/_ref/refwsharness): per-turn CMs, monitor statistics, ion-cloud sums,
    bucket z-centroids, final slices, and the first 64 generated macro-electrons."""
    from variants import write_ws_variant
    with tempfile.TemporaryDirectory() as d:
        conf, mode = write_ws_variant(name, d)
        text = open(conf).read().replace(os.path.abspath(os.path.join(d, "fill_ws.txt")), "@FILL@")
        r = subprocess.run([WS_HARNESS, conf, d, "0", str(mode)], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(r.stderr)
        ld = lambda f: np.loadtxt(os.path.join(d, f), ndmin=2)  # noqa: E731
        dist = np.concatenate([ld(f)[:64, 1:] for f in ("Ldist", "Hdist", "Vdist")], axis=1)
        out = os.path.join(HERE, name + ".npz")
        np.savez_compressed(out, conf=np.array(text), mode=mode, cm=ld("cm.txt"), mon=ld("mon.txt"),
                            ion=ld("ion.txt"), zcm=ld("zcm.txt"), slices=ld("slices.txt"), dist=dist)
    return out


def make_wofz():
    """w(z) of the reference's complex_errorfunc_wofz_algo680 on 1,082 points."""
    r = subprocess.run([WOFZ_DUMP], capture_output=True, text=True, check=True)
    a = np.array([[float(t) for t in l.split()] for l in r.stdout.splitlines()])
    out = os.path.join(HERE, "wofz_ref.npz")
    np.savez_compressed(out, x=a[:, 0], y=a[:, 1], u=a[:, 2], v=a[:, 3], flag=a[:, 4].astype(int))
    return out


if __name__ == "__main__":
    from variants import WS_VARIANTS
    names = sys.argv[1:] or (sorted(VARIANTS) + ["e2e:" + c for c in E2E_CASES] + sorted(WS_VARIANTS)
                             + ["wofz", "inp"])
    for n in names:
        if n == "inp":
            print(n, make_inp())
        elif n.startswith("e2e:"):
            print(n, make_e2e(n[4:]))
        elif n in WS_VARIANTS:
            print(n, make_ws(n))
        elif n == "wofz":
            print(n, make_wofz())
        else:
            p = make(n)
            print(n, os.path.getsize(p))
