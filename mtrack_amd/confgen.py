"""Writers of mbtrack .conf inputs for tests and benchmarks.

`ring_conf(**overrides)` renders the reference's INI format (sections
[tracking] [ring] [distribution] [selffield] [resonator_N] [harmonic_cavity_N]
[longrange_resonator_N] [rf_feedback] [scan]), keys as read by the reference's
src/input.c:278-850.  The defaults describe the 3 GeV, 354 m, h = 416 ring of
the benchmark workloads: Qx 18.2, Qy 10.3, Qs 0.006 (Vrf0 5.53 MV at
alpha_c 3e-4, U0 1 MeV), chromaticity 1/1, damping 10k/10k/5k turns,
6-D Gaussian sx 100 um, sy 10 um, sz 3 mm, sdelta 1e-3.
"""
import os

BASE = {
    "tracking": {
        "jobtitle": "mtrack-mi355x test ring",
        "filling": "UNIFORM",
        "BunchModel": "WEAK",
        "TrackPlane_LON": 1, "TrackPlane_HOR": 1, "TrackPlane_VER": 1,
        "NrevTot": 100, "NrevMon": 10,
        "EnableRW_short": 0, "EnableRW_long": 0, "EnableFBII": 0,
        "EnableQuantum": 0, "EnableIdealHC": 0, "EnableAmpinv_out": 1,
    },
    "ring": {
        "E0": 3.0, "Lc": 354.0, "rho0": 8.0,
        "betaH": 10.0, "betaV": 10.0, "dispH": 0.0, "alphaH": 0.0, "alphaV": 0.0, "disppH": 0.0,
        "emittanceH": 1.0, "couplbeta": 0.01,
        "ac": 3.0e-4, "Je": 2.0, "Vrf0": 5.53, "frf": 352.3, "Iring": 208.0,
        "ApertureH": 20.0, "ApertureV": 10.0,
        "QH0": 18.2, "QV0": 10.3, "Cx": 1.0, "Cz": 1.0,
        "U0": 1000.0, "taue": 5.904,
    },
    "distribution": {
        "Nslc": 0, "Np": 1000,
        "nGen_LON": 1, "nGen_HOR": 1, "nGen_VER": 1,
        "iseed_LON": 1001, "iseed_HOR": 2002, "iseed_VER": 3003,
        "gendst_datafile_LON": "Ldist", "gendst_datafile_HOR": "Hdist", "gendst_datafile_VER": "Vdist",
        "xtauCM_offset": 0.0, "xepsCM_offset": 0.0, "sgm_xtau": 0.010007, "sgm_xeps": 1.0e-3,
        "xxCM_offset": 0.01, "xpCM_offset": 0.0, "sgm_xx": 0.1, "sgm_xp": 0.01,
        "zzCM_offset": 0.001, "zpCM_offset": 0.0, "sgm_zz": 0.01, "sgm_zp": 0.001,
    },
}

ORDER = ["tracking", "ring", "selffield", "distribution"]


def ring_conf(sections=None, extra_sections=(), **overrides):
    """Return .conf text.  `overrides` maps "section.key" (dots replaced by __)
    to a value; `sections` adds/overrides whole sections; `extra_sections` is a
    list of (name, dict) appended in order (resonators, cavities, ...)."""
    secs = {k: dict(v) for k, v in BASE.items()}
    for name, d in (sections or {}).items():
        secs.setdefault(name, {}).update(d)
    for key, val in overrides.items():
        sec, k = key.split("__", 1)
        secs.setdefault(sec, {})[k] = val
    out = []
    names = [n for n in ORDER if n in secs] + [n for n in secs if n not in ORDER]
    for name in names:
        out.append("[%s]" % name)
        for k, v in secs[name].items():
            out.append("  %s = %s" % (k, _fmt(v)))
        out.append("")
    for name, d in extra_sections:
        out.append("[%s]" % name)
        for k, v in d.items():
            out.append("  %s = %s" % (k, _fmt(v)))
        out.append("")
    return "\n".join(out)


def _fmt(v):
    if isinstance(v, float):
        return repr(v)
    return str(v)


def write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return path


def filling_file(path, buckets, ratio=1.0, out=(0,)):
    """Filling file lines `bucket ratio outflag` (reference src/input.c:1461)."""
    with open(path, "w") as f:
        for b in buckets:
            f.write("%d %g %d\n" % (b, ratio, 1 if b in out else 0))
    return path


def small_ring(h_target=20):
    """Lc giving harmonic number h_target at 352.3 MHz."""
    return h_target * 2.997925e8 / 352.3e6


# ---- named workloads -------------------------------------------------------

def wl_production(turns=20000, np_=100000, rw=True, bbr=True):
    """416 bunches x np_ macroparticles, 100 slices, broadband resonator
    (10 GHz, Q 1, 5 kOhm) + resistive wall over 50 past turns, 0.5 mA/bunch,
    quantum excitation off."""
    extra = []
    if bbr:
        extra.append(("resonator_1", {"Rs": 5000.0, "fres": 10.0, "Qfactor": 1.0, "plane": "LON"}))
    return ring_conf(
        sections={
            "selffield": {"Ncell": 100, "Nsigma": 5.0},
            "tracking": {"NrevTot": turns, "NrevMon": 100, "EnableRW_long": 1 if rw else 0,
                         "Nmlt": 50, "EnableQuantum": 0},
            "ring": {"Iring": 208.0, "rhorw": 1.7e-8, "beffL1": 0.012, "beffL2": 0.012,
                     "beffH3": 0.012, "beffV3": 0.012},
            "distribution": {"Np": np_},
        },
        extra_sections=extra)


# ---- batched ensemble of independent rings ------------------------------------

def wl_ensemble(d, nrings=256, seed=20240607, turns=10000, np_range=(10000, 100000), nb_range=(32, 416),
                ib_range=(0.1, 5.0), chrom_range=(-1.0, 3.0)):
    """Writes nrings independent ring configurations into directory d and
    returns their paths.  Each ring: the benchmark machine (h = 416), a bunch
    count and macroparticle count drawn log-uniform from nb_range / np_range,
    a bunch current uniform in ib_range [mA] and horizontal and vertical
    chromaticities uniform in chrom_range, equally spaced bunches, 100 slices,
    broadband resonator (10 GHz, Q 1, 5 kOhm), resistive wall over 50 past
    turns, the main cavity as a long-range resonator with RF beam-loading
    feedback, quantum excitation off.  Draws come from numpy's PCG64 with
    the given seed."""
    import numpy as np

    rng = np.random.default_rng(seed)
    os.makedirs(d, exist_ok=True)
    paths = []
    h = 416
    for k in range(nrings):
        nb = int(round(np.exp(rng.uniform(np.log(nb_range[0]), np.log(nb_range[1])))))
        npb = int(round(np.exp(rng.uniform(np.log(np_range[0]), np.log(np_range[1])))))
        ib = float(rng.uniform(*ib_range))
        cx, cz = (float(v) for v in rng.uniform(chrom_range[0], chrom_range[1], 2))
        nb = max(1, min(nb, h))
        fill = os.path.join(d, "fill_%03d.txt" % k)
        filling_file(fill, [int(i * h / nb) for i in range(nb)], out=())
        text = ring_conf(
            sections={
                "selffield": {"Ncell": 100, "Nsigma": 5.0},
                "tracking": {"NrevTot": turns, "NrevMon": 100, "EnableRW_long": 1, "Nmlt": 50,
                             "EnableQuantum": 0, "filling": fill},
                "ring": {"Iring": ib * nb, "Cx": cx, "Cz": cz, "rhorw": 1.7e-8, "beffL1": 0.012,
                         "beffL2": 0.012, "beffH3": 0.012, "beffV3": 0.012},
                "distribution": {"Np": npb, "iseed_LON": 1001 + 3 * k, "iseed_HOR": 1002 + 3 * k,
                                 "iseed_VER": 1003 + 3 * k},
            },
            extra_sections=[("resonator_1", {"Rs": 5000.0, "fres": 10.0, "Qfactor": 1.0, "plane": "LON"}),
                            ("harmonic_cavity_1", {"m": 1, "Rs": 1e6, "detune": -40e3, "Qfactor": 4e4}),
                            ("rf_feedback", {"resonator": 1, "averaging_length": 416})])
        paths.append(write(os.path.join(d, "ring_%03d.conf" % k), text))
    return paths


def wl_c1(d, turns=2000, np_=1000, iring=208.0, beff=0.012, ncell=50):
    """The small multi-bunch ring: 64 bunches in buckets int(6.5 k) of h = 416,
    np_ macroparticles each, quantum excitation on (damping 10k / 10k / 5k
    turns), chromaticity 1 / 1, copper resistive wall (rho 1.7e-8 Ohm m, radius
    beff) over Nmlt = 19 past turns, a 10 GHz broadband resonator in Ncell
    slices, Iring [mA] shared equally.  Writes the filling file and the .conf
    into d and returns the .conf path."""
    os.makedirs(d, exist_ok=True)
    fill = os.path.join(d, "fill_c1.txt")
    filling_file(fill, [int(k * 6.5) for k in range(64)], out=(0,))
    text = ring_conf(
        sections={
            "selffield": {"Ncell": ncell, "Nsigma": 5.0},
            "tracking": {"NrevTot": turns, "NrevMon": 10, "EnableRW_long": 1, "Nmlt": 19, "EnableQuantum": 1,
                         "filling": fill},
            "ring": {"Iring": iring, "rhorw": 1.7e-8, "beffL1": beff, "beffL2": beff, "beffH3": beff,
                     "beffV3": beff},
            "distribution": {"Np": np_},
        },
        extra_sections=[("resonator_1", {"Rs": 5000.0, "fres": 10.0, "Qfactor": 1.0, "plane": "LON"})])
    return write(os.path.join(d, "c1.conf"), text)
