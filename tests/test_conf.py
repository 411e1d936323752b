"""The .conf grammar, pinned by the reference's own test (src/test_input.c with
its data file src/test_input.input, kept as tests/golden/test_input.input)."""
import ctypes
import os

import pytest

import mtrack_amd as m

INPUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_input.input")


def get_d(sec, key):
    v = ctypes.c_double(-1.0)
    L = m.lib()
    L.mbt_conf_get_double.argtypes = [ctypes.c_char_p] * 3 + [ctypes.POINTER(ctypes.c_double)]
    ok = L.mbt_conf_get_double(INPUT.encode(), sec.encode(), key.encode(), ctypes.byref(v))
    return ok, v.value


def get_i(sec, key):
    v = ctypes.c_int(-7)
    L = m.lib()
    L.mbt_conf_get_int.argtypes = [ctypes.c_char_p] * 3 + [ctypes.POINTER(ctypes.c_int)]
    ok = L.mbt_conf_get_int(INPUT.encode(), sec.encode(), key.encode(), ctypes.byref(v))
    return ok, v.value


def get_s(sec, key):
    buf = ctypes.create_string_buffer(256)
    L = m.lib()
    L.mbt_conf_get_str.argtypes = [ctypes.c_char_p] * 4 + [ctypes.c_int]
    ok = L.mbt_conf_get_str(INPUT.encode(), sec.encode(), key.encode(), buf, 256)
    return ok, buf.value.decode()


def have(sec):
    L = m.lib()
    L.mbt_conf_have_section.argtypes = [ctypes.c_char_p] * 2
    return L.mbt_conf_have_section(INPUT.encode(), sec.encode())


def test_bogus_lines_rejected():
    assert get_d("", "bogus1")[0] == 0
    for k in ("bogus1", "bogus2", "bogus3"):
        assert get_d("bogus_land", k)[0] == 0
    assert get_i("bogus_land", "bogus4")[0] == 0


def test_comments_and_sections():
    assert get_d("general", "cx")[0] == 0
    assert get_d("physic", "a")[0] == 0
    assert get_d("general", "clight")[0] == 0
    assert have("general") and have("physic")
    assert not have("ghost")
    assert not have("twin")  # sections must be unique
    ok, v = get_d("general", "DoNotExist")
    assert ok == 0 and v == -1.0  # value untouched


@pytest.mark.parametrize("key,val", [("a", 1.1), ("b", 2.2), ("c", 3.3), ("d", 4.4), ("e", 5.5), ("f", 6.6),
                                     ("g", 7.7), ("h", 8.8), ("fortytwo", 42.0)])
def test_doubles(key, val):
    assert get_d("general", key) == (1, val)


def test_other_values():
    assert get_d("physic", "clight") == (1, 2.997925e08)
    assert get_s("general", "message") == (1, "hello World!")
    assert get_s("general", "messagewident") == (1, "hello World!")
    assert get_i("general", "n") == (1, -2)
    assert get_i("general", "m") == (1, 5)


def test_unsupported_and_missing(tmp_path):
    with pytest.raises(m.MbtError):
        m.Config(str(tmp_path / "nothere.conf"))
    p = tmp_path / "x.inp"
    p.write_text("legacy\n")
    with pytest.raises(m.MbtError):
        m.Config(str(p))
    p = tmp_path / "bad.conf"
    p.write_text("[tracking]\n jobtitle = t\n BunchModel = WEAK\n TrackPlane_LON = 0\n")
    with pytest.raises(m.MbtError):
        m.Config(str(p))


def test_legacy_inp_params_bitwise(tmp_path):
    """The legacy fixed-order input (read_input_file, src/input.c:198-276, comment
    lines starting with C/c): every derived parameter equals the reference's,
    from its own read_input + setup_* on the same file (tests/golden/inp_params.npz,
    made by oracle/_ref/refharness; the harness and this reader both set
    NrevScan = NrevTot + 1, which the reference's .inp path leaves 0).  The
    reference's .inp path never sets U0 and taue, so Urad is 0 and the damping
    constants are infinite there and here alike."""
    import numpy as np
    from variants import write_inp
    g = np.load(os.path.join(os.path.dirname(INPUT), "inp_params.npz"))
    cfg = m.Config(write_inp(str(tmp_path)), str(tmp_path) + "/")
    dump = str(tmp_path / "mine.txt")
    cfg.dump(dump)
    mine, outs = {}, []
    for line in open(dump):
        p = line.split()
        if len(p) == 2 and p[0] == "bunch_out":
            outs.append(int(p[1]))
        elif len(p) == 2:
            try:
                mine[p[0]] = float(p[1])
            except ValueError:
                pass
    checked = 0
    for k, v in zip(g["param_names"], g["param_values"]):
        k = str(k)
        if k == "bunch_out" or k not in mine:
            continue
        assert mine[k] == v or (np.isnan(v) and np.isnan(mine[k])), (k, mine[k], v)
        checked += 1
    assert checked > 60
    assert outs == [int(v) for k, v in zip(g["param_names"], g["param_values"]) if str(k) == "bunch_out"]
    assert cfg.get("Gzix") == 1.0 and cfg.geti("Np") == 1000 and cfg.geti("Nmlt") == 19
