"""Token-wise comparison of the reference's text outputs with ours: words must be
identical; numbers may differ by at most `units` in their last printed digit
(the GPU path agrees with the reference to ~1e-15 relative, so a printed digit
flips only when a value lies on a rounding boundary)."""
import re

NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?$")


def last_digit_unit(tok):
    m = re.match(r"^[-+]?(\d*)\.?(\d*)(?:[eE]([-+]?\d+))?$", tok)
    frac = m.group(2) or ""
    exp = int(m.group(3) or 0)
    return 10.0 ** (exp - len(frac))


def compare(a_text, b_text, units=2.0, a_nan_matches_number=False):
    """Returns a list of mismatch descriptions (empty when equivalent).  With
    a_nan_matches_number, a nan token of the first text (the reference's
    output) stands for any number of the second."""
    out = []
    al, bl = a_text.splitlines(), b_text.splitlines()
    if len(al) != len(bl):
        return ["line count %d != %d" % (len(al), len(bl))]
    for i, (x, y) in enumerate(zip(al, bl)):
        xt, yt = x.split(), y.split()
        if len(xt) != len(yt):
            out.append("line %d: token count %d != %d" % (i + 1, len(xt), len(yt)))
            continue
        for p, q in zip(xt, yt):
            if p == q:
                continue
            if a_nan_matches_number and p.lstrip("+-").lower() == "nan" and NUM.match(q):
                continue
            if NUM.match(p) and NUM.match(q):
                tol = units * max(last_digit_unit(p), last_digit_unit(q))
                if abs(float(p) - float(q)) <= tol * (1 + 1e-9):
                    continue
            out.append("line %d: %r != %r" % (i + 1, p, q))
    return out
