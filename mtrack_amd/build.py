"""Builds libmbtrack_mi355x.so (host C + gfx950 HIP kernels) and the
mbtrack-mi355x command-line program in-tree.

Host C units are compiled by gcc with -ffp-contract=off so the derived machine
parameters round exactly like the reference's x86-64 build; HIP units by hipcc
for gfx950 only.  Usage: python -m mtrack_amd.build  (or build() from Python).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmbtrack_mi355x.so")
EXE = os.path.join(HERE, "mbtrack-mi355x")

HOST_C = ["conf.c", "bunchgen.c", "writers.c", "driver.c", "ws_driver.c"]
HIP_SRC = ["ring.hip", "wofz.hip", "ws.hip"]
ARCH = "gfx950"

CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-I" + INC, "-I" + CSRC]
# -amdgpu-mfma-vgpr-form: MFMA accumulators in VGPRs (gfx950 allows both); the
# default heuristic kept the RW GEMM's loop-carried accumulators in VGPRs and
# copied them to and from AGPRs around every MFMA
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form=1",
            "-munsafe-fp-atomics", "-I" + INC, "-I" + CSRC, "-Wno-unused-result"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in HOST_C:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        o = os.path.join(OBJ, src + ".o")
        _run(["gcc"] + CFLAGS + ["-c", path, "-o", o])
        objs.append(o)
    hipcc = _hipcc()
    for src in HIP_SRC:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        o = os.path.join(OBJ, src + ".o")
        _run([hipcc] + HIPFLAGS + ["-c", path, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lm", "-lpthread"])
    os.replace(tmp, LIB)
    main = os.path.join(CSRC, "main.c")
    if os.path.exists(main):
        _run(["gcc"] + CFLAGS + ["-o", EXE, main, "-L" + HERE, "-lmbtrack_mi355x",
                                 "-Wl,-rpath,$ORIGIN", "-lm"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
