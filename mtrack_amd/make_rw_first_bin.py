#!/usr/bin/env python3
"""Writes mtrack_amd/csrc/rw_first_bin_data.h: the first-bin values of the
short-range resistive-wall wake that the reference tabulates against
u = bin width / s0 (src/rw_table.c:5-306, columns u, longitudinal,
transverse).  The reference gives no closed form for them, so they are data
that has to be alike; this script copies the numbers (as printed, so every
double is the same) into three arrays of the build's own header.

usage: python3 mtrack_amd/make_rw_first_bin.py [/root/reference/src/rw_table.c]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/rw_table.c"
    text = open(src).read()
    body = text[text.index("{") + 1:text.index("}")]
    toks = re.findall(r"[-+]?\d+\.\d+e[-+]\d+", body)
    assert len(toks) % 3 == 0
    rows = [toks[i:i + 3] for i in range(0, len(toks), 3)]
    out = ["/* First-bin values of the short-range resistive-wall wake against",
           " * u = bin width / s0, as tabulated by the reference (src/rw_table.c:5-306);",
           " * written by mtrack_amd/make_rw_first_bin.py. */",
           "#define RWFB_N %d" % len(rows)]
    for name, col in (("rwfb_u", 0), ("rwfb_lon", 1), ("rwfb_trans", 2)):
        out.append("static const double %s[RWFB_N] = {" % name)
        vals = [r[col] for r in rows]
        for i in range(0, len(vals), 6):
            out.append("    " + ", ".join(vals[i:i + 6]) + ",")
        out.append("};")
    path = os.path.join(HERE, "csrc", "rw_first_bin_data.h")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    print(path, len(rows), "rows")


if __name__ == "__main__":
    main()
