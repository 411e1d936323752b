"""Small seeded ring configurations shared by the parity tests and the golden
fixture generator.  Each exercises one part of the reference's weak-weak turn."""
import os

from mtrack_amd import confgen as cg

L20 = cg.small_ring(20)
RW = {"rhorw": 1.7e-8, "beffL1": 0.012, "beffL2": 0.012, "beffH3": 0.012, "beffV3": 0.012}
BBR_L = ("resonator_1", {"Rs": 5000.0, "fres": 10.0, "Qfactor": 1.0, "plane": "LON"})


def _ring(**kw):
    d = {"Lc": L20, "Iring": 40.0}
    d.update(RW)
    d.update(kw)
    return d


VARIANTS = {
    # broadband resonators in all planes + RW long-range, quantum excitation off
    "wakes": dict(sections={"selffield": {"Ncell": 50, "Nsigma": 5.0},
                            "tracking": {"NrevTot": 60, "NrevMon": 10, "EnableRW_long": 1, "Nmlt": 3},
                            "ring": _ring(), "distribution": {"Np": 1000}},
                  extra_sections=[BBR_L,
                                  ("resonator_2", {"Rs": 2e5, "fres": 8.0, "Qfactor": 1.0, "plane": "VER"}),
                                  ("resonator_3", {"Rs": 2e5, "fres": 8.0, "Qfactor": 1.0, "plane": "HOR"})]),
    # quantum excitation and radiation damping
    "qe": dict(sections={"tracking": {"NrevTot": 40, "EnableQuantum": 1}, "ring": _ring(),
                         "distribution": {"Np": 500}}),
    # harmonic cavities (main + 3rd) as long-range resonators with RF feedback
    "lr_fb": dict(sections={"selffield": {"Ncell": 40, "Nsigma": 5.0}, "tracking": {"NrevTot": 30},
                            "ring": _ring(Iring=200.0), "distribution": {"Np": 400}},
                  extra_sections=[("harmonic_cavity_1", {"m": 1, "Rs": 5e6, "detune": -40e3, "Qfactor": 4e4}),
                                  ("harmonic_cavity_2", {"m": 3, "Rs": 2e6, "detune": 50e3, "Qfactor": 2e4}),
                                  ("rf_feedback", {"resonator": 1, "averaging_length": 30})]),
    # transverse long-range resonators (HOMs) + LON broadband, QE on
    "lr_hv": dict(sections={"selffield": {"Ncell": 40, "Nsigma": 5.0},
                            "tracking": {"NrevTot": 30, "EnableQuantum": 1},
                            "ring": _ring(), "distribution": {"Np": 400}},
                  extra_sections=[("longrange_resonator_1", {"plane": "HOR", "fres": 1.2, "Rs": 1e6, "Qfactor": 3e3}),
                                  ("longrange_resonator_2", {"plane": "VER", "fres": 1.3, "Rs": 1e6, "Qfactor": 3e3}),
                                  BBR_L]),
    # active harmonic cavity, amplitude-dependent tunes, optics with alpha and dispersion
    "ahc_amp": dict(sections={"tracking": {"NrevTot": 30, "EnableQuantum": 1},
                              "ring": _ring(CHH=100.0, CHV=50.0, CVH=30.0, CVV=70.0, alphaH=0.5, alphaV=-0.3,
                                            dispH=0.1, disppH=0.01),
                              "distribution": {"Np": 300}},
                    extra_sections=[("active_HC_1", {"Vpeak": 3e5, "nHC": 3.0, "phi_aHC": 0.3})]),
    # ideal harmonic cavity
    "ideal": dict(sections={"tracking": {"NrevTot": 30, "EnableIdealHC": 1, "EnableQuantum": 1},
                            "ring": _ring(mIdeal=3), "distribution": {"Np": 300}}),
    # 64 bunches in 416 buckets from a filling file, RW trigger, automatic Nmlt
    "fromfile": dict(sections={"selffield": {"Ncell": 60, "Nsigma": 6.0},
                               "tracking": {"NrevTot": 40, "filling": "@FILL@", "EnableRW_long": 1, "Nmlt": 0,
                                            "triggerRW": 1},
                               "ring": dict(Iring=30.0, **RW), "distribution": {"Np": 200}},
                     extra_sections=[BBR_L]),
    # current scan (scan_type 1)
    "scan": dict(sections={"selffield": {"Ncell": 40, "Nsigma": 5.0},
                           "tracking": {"NrevTot": 30, "EnableRW_long": 1, "Nmlt": 2},
                           "ring": _ring(), "distribution": {"Np": 200}},
                 extra_sections=[BBR_L, ("scan", {"scan_type": 1, "Nscan": 3, "NrevScan": 10, "scan_start": 20.0,
                                                  "scan_step": 10.0})]),
    # resistive + inductive smoothing terms (bunch well inside the binning range)
    "rsis": dict(sections={"selffield": {"Ncell": 60, "Nsigma": 9.0, "RsisL": 300.0, "aindL": 5.0},
                           "tracking": {"NrevTot": 30}, "ring": _ring(), "distribution": {"Np": 400}},
                 extra_sections=[BBR_L]),
    # short-range resistive wall in all planes (tabulated first bins, 1/t^(3/2)
    # and 1/sqrt(t) tails) with the long-range RW; bin width 26 s0
    "rw_short": dict(sections={"selffield": {"Ncell": 50, "Nsigma": 5.0},
                               "tracking": {"NrevTot": 30, "EnableRW_short": 1, "EnableRW_short_LON": 1,
                                            "EnableRW_long": 1, "Nmlt": 3},
                               "ring": _ring(Iring=60.0, beffH4=0.012, beffV4=0.012),
                               "distribution": {"Np": 400}}),
    # edge: one bunch, an odd macroparticle count below one work-item chunk, six
    # slices over +-1.5 sigma (particles in both trash counters), RW over one turn
    "edge_single": dict(sections={"selffield": {"Ncell": 6, "Nsigma": 1.5},
                                  "tracking": {"NrevTot": 25, "filling": "SINGLE", "EnableRW_long": 1, "Nmlt": 1},
                                  "ring": _ring(Iring=5.0), "distribution": {"Np": 37}},
                        extra_sections=[BBR_L]),
    # edge: every bucket filled with an odd count, bins over +-1 sigma only (most
    # particles clamped into the edge bins by the selffield rule and dropped by
    # the long-range rule), a long-range resonator reading those histograms
    "edge_trash": dict(sections={"selffield": {"Ncell": 10, "Nsigma": 1.0},
                                 "tracking": {"NrevTot": 25},
                                 "ring": _ring(), "distribution": {"Np": 999}},
                       extra_sections=[BBR_L, ("longrange_resonator_1", {"plane": "LON", "fres": 352.5, "Rs": 1e6,
                                                                         "Qfactor": 3e3})]),
    # edge: one macroparticle per bunch (the generator's offset-only case, zero
    # rms in every plane), 20 slices, RW over two turns
    "edge_np1": dict(sections={"selffield": {"Ncell": 20, "Nsigma": 5.0},
                               "tracking": {"NrevTot": 20, "EnableRW_long": 1, "Nmlt": 2},
                               "ring": _ring(), "distribution": {"Np": 1}},
                     extra_sections=[BBR_L]),
    # wake functions read from files (construct_greensfunc_file): a longitudinal
    # file in the LON plane, a transverse one in VER and HOR, each sampled finely,
    # then coarsely (steps of about 2.5 bins, the branch that fills several bins
    # from one pair of samples), then finely again past the last bin; a broadband
    # resonator adds to the longitudinal function
    "wakefile": dict(sections={"selffield": {"Ncell": 50, "Nsigma": 5.0},
                               "tracking": {"NrevTot": 30}, "ring": _ring(Iring=60.0),
                               "distribution": {"Np": 400}},
                     extra_sections=[BBR_L,
                                     ("wake_file_1", {"filename": "@WAKE_L@", "plane": "LON"}),
                                     ("wake_file_2", {"filename": "@WAKE_T@", "plane": "VER"}),
                                     ("wake_file_3", {"filename": "@WAKE_T@", "plane": "HOR"})]),
}

FILL_BUCKETS = [int(k * 6.5) for k in range(64)]
WAKE_FILES = {"@WAKE_L@": "wake_lon.dat", "@WAKE_T@": "wake_trans.dat"}


def wake_samples():
    """Sample times [s] of the wake files (the reader compares them with the bin
    edges k dTau in seconds): two negative times (skipped by the reader), 0 to
    30 ps in steps of 0.71 ps, 30 to 72 ps in steps of 5.1 ps, then to 125 ps in
    steps of 1.3 ps (bins of 2.0422 ps: 50 slices over +-5 sigma of
    10.007 ps)."""
    import numpy as np
    return 1e-12 * np.concatenate([[-4.0, -1.0], np.arange(0.0, 30.0, 0.71), np.arange(30.0, 72.0, 5.1),
                                   np.arange(72.0, 125.0, 1.3)])


def write_wake_files(d):
    """The two wake files of the "wakefile" variant in d: a damped cosine
    (longitudinal, amplitude 1e14) and a damped sine (transverse, 3e15), in the
    units of the Green functions they are added to."""
    import numpy as np
    t = wake_samples()
    out = {}
    for key, (name, amp, shape) in {"@WAKE_L@": ("wake_lon.dat", 1.0e14, np.cos),
                                    "@WAKE_T@": ("wake_trans.dat", 3.0e15, np.sin)}.items():
        w = amp * np.exp(-np.maximum(t, 0.0) / 20e-12) * shape(2 * np.pi * t / 30e-12)
        path = os.path.join(d, name)
        with open(path, "w") as f:
            for a, b in zip(t, w):
                f.write("%.10g %.10g\n" % (a, b))
        out[key] = os.path.abspath(path)
    return out


# end-to-end cases beyond the variants themselves: (variant, section overrides)
E2E_VARIANTS = {
    # 600 turns with potentials only from turn 500: the drop-in program tracks
    # 240-turn blocks first, then blocks split at every potentials output
    "wakes_blocks": ("wakes", {"tracking": {"NrevTot": 600, "NrevMon": 20, "NrevPotentialsOut": 500}}),
    # a monitor interval longer than the library's 256-turn statistics block
    "wakes_mon300": ("wakes", {"tracking": {"NrevTot": 700, "NrevMon": 300, "NrevPotentialsOut": 600}}),
}


def write_e2e_variant(name, d):
    """A variant or an E2E_VARIANTS case written into d; returns the .conf path."""
    if name in E2E_VARIANTS:
        base, over = E2E_VARIANTS[name]
        return write_variant(base, d, **over)
    return write_variant(name, d)


def write_variant(name, d, **section_overrides):
    """Writes the variant's .conf (and filling file) into directory d; returns the
    path.  section_overrides maps a section name to keys that replace the
    variant's own (e.g. selffield={"Ncell": 40})."""
    os.makedirs(d, exist_ok=True)
    fill = os.path.join(d, "fill.txt")
    cg.filling_file(fill, FILL_BUCKETS, out=(0, 13))
    v = dict(VARIANTS[name])
    secs = {k: dict(x) for k, x in v.get("sections", {}).items()}
    for sec, keys in section_overrides.items():
        secs.setdefault(sec, {}).update(keys)
    v["sections"] = secs
    text = cg.ring_conf(**v).replace("@FILL@", os.path.abspath(fill))
    if "@WAKE_" in text:
        for key, path in write_wake_files(d).items():
            text = text.replace(key, path)
    return cg.write(os.path.join(d, name + ".conf"), text)


def conf_template(text, d):
    """A written variant's .conf text with the paths of its filling and wake files
    in d replaced by their placeholders (the form the golden fixtures store)."""
    text = text.replace(os.path.abspath(os.path.join(d, "fill.txt")), "@FILL@")
    for key, name in WAKE_FILES.items():
        text = text.replace(os.path.abspath(os.path.join(d, name)), key)
    return text


# ---- weak-strong (BunchModel = STRONG) configurations: slice-centroid bunches,
# fast beam-ion interaction, resistive-wall long-range on the centroid history
WS_FILL = {20: [int(k * 20 / 12) for k in range(12)], 416: [int(k * 416 / 12) for k in range(12)]}


def _ws(h, fbii, rw, nmlt=3, turns=30, np_=2000, filling="@FILL@", nslc=20):
    return dict(sections={
        "tracking": {"BunchModel": "STRONG", "EnableFBII": fbii, "NrevTot": turns, "NrevMon": 10,
                     "EnableRW_long": rw, "Nmlt": nmlt, "EnableQuantum": 0, "filling": filling},
        "ring": dict(Lc=cg.small_ring(h), Iring=100.0, **RW),
        "distribution": {"Np": np_, "Nslc": nslc}})


# name: (configuration, bunch-current mode: 1 = currents of the filling, 0 = the
# reference's STRONG-mode zeros)
WS_VARIANTS = {
    "ws_fbii": (_ws(20, 1, 1), 1),
    "ws_fbii_refcur": (_ws(20, 1, 1), 0),
    "ws_optics": (_ws(20, 0, 0), 1),
    "ws_h416": (_ws(416, 1, 1, nmlt=12), 1),
    # 19 slices a side: 39 centroids (the reference bins at 0.3 sigma from
    # -7 sigma, so it needs about 17 or more), 39 = 9 x 4 + 3 slices
    "ws_s39": (_ws(20, 1, 1, nslc=19), 1),
    # every bucket filled (the reference program's manager addresses bucket kb
    # as rank kb + 1, so its end-to-end run needs buckets 0..Nbunch-1 filled)
    "ws_uniform": (_ws(20, 1, 1, filling="UNIFORM"), 0),
}


def write_ws_variant(name, d):
    """Writes a weak-strong variant's .conf and filling file into d; returns
    (path, bunch-current mode)."""
    v, mode = WS_VARIANTS[name]
    os.makedirs(d, exist_ok=True)
    h = round(v["sections"]["ring"]["Lc"] * 352.3e6 / 2.997925e8)
    fill = os.path.join(d, "fill_ws.txt")
    cg.filling_file(fill, WS_FILL[h], out=())
    text = cg.ring_conf(**v).replace("@FILL@", os.path.abspath(fill))
    return cg.write(os.path.join(d, name + ".conf"), text), mode


# ---- quantum-excitation statistics: 8 bunches in h = 8, every initial sigma
# twice its nominal value, strong damping (taue = 100 turns, U0 = 4 MeV: about
# 750 turns of transverse amplitude damping through the RF focusing term), so
# every plane relaxes towards its equilibrium within the run
QE_STAT_TURNS = 1200


def write_qe_stat(d, np_=4000, turns=QE_STAT_TURNS, seed=0, **ring):
    os.makedirs(d, exist_ok=True)
    T0 = 354.0 / 2.997925e8
    r = {"Lc": cg.small_ring(8), "taue": 100 * T0 * 1e3, "Iring": 1.0, "U0": 4000.0}
    r.update(ring)
    text = cg.ring_conf(sections={
        "tracking": {"NrevTot": turns, "NrevMon": 10, "EnableQuantum": 1},
        "ring": r,
        "distribution": {"Np": np_, "sgm_xtau": 0.02, "sgm_xeps": 2.0e-3, "sgm_xx": 0.2, "sgm_xp": 0.02,
                         "sgm_zz": 0.02, "sgm_zp": 0.002, "xxCM_offset": 0.0, "zzCM_offset": 0.0,
                         "iseed_LON": 11 + seed, "iseed_HOR": 22 + seed, "iseed_VER": 33 + seed}})
    return cg.write(os.path.join(d, "qe_stat.conf"), text)


def sigma2_curves(stats):
    """Bunch-averaged rms^2 per turn and plane, [turns][3], from stats [turns][G][12]."""
    import numpy as np
    return (np.asarray(stats)[..., 3:6] ** 2).mean(axis=1)


def qe_stat_compare(a, b):
    """(relative difference of the last-third means, max per-turn relative
    difference), each per plane, of two sigma2_curves outputs."""
    import numpy as np
    w = slice(-(len(a) // 3), None)
    eq = np.abs(a[w].mean(0) - b[w].mean(0)) / a[w].mean(0)
    return eq, (np.abs(a - b) / a).max(0)


# ---- the legacy fixed-order input (read_input_file, src/input.c:198-276): the
# benchmark ring's values in its 21 data lines, with comment lines between
INP_LINES = [
    ("C job title", "mtrack-mi355x legacy input"),
    ("C E0 [GeV]", "3.0"),
    ("C Lc [m]  rho0 [m]", "354.0 8.0"),
    ("C betaH betaV dispH", "10.0 10.0 0.0"),
    ("C alphaH alphaV disppH", "0.0 0.0 0.0"),
    ("C emittanceH couplbeta", "1.0 0.01"),
    ("C ac Je", "3.0e-4 2.0"),
    ("C Vrf0 [MV] frf [MHz]", "5.53 352.3"),
    ("C Iring [mA]", "208.0"),
    ("C ApertureH ApertureV", "20.0 10.0"),
    ("C chromaticity H V", "1.0 1.0"),
    ("C TrackPlane LON HOR VER", "1 1 1"),
    ("C NrevTot Nmlt NrevMon", "100 19 10"),
    ("C Np", "1000"),
    ("C nGen iseed per plane", "1 1001 1 2002 1 3003"),
    ("C distribution files", "Ldist"),
    ("c", "Hdist"),
    ("c", "Vdist"),
    ("C tau: CM offset, eps: CM offset, sgm_tau, sgm_eps", "0.0 0.0 0.010007 1.0e-3"),
    ("C x", "0.01 0.0 0.1 0.01"),
    ("C z", "0.001 0.0 0.01 0.001"),
]


def write_inp(d, name="legacy"):
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, name + ".inp")
    with open(path, "w") as f:
        for comment, data in INP_LINES:
            f.write(comment + "\n" + data + "\n")
    return path


# ---- growth rates --------------------------------------------------------------
# A uniformly filled ring with the resistive-wall long-range wake alone (no
# short-range wake, chromaticity 0, quantum excitation and with it radiation
# damping off), every particle of a bunch on the same point: rigid bunches whose
# centroids follow the coupled-bunch modes exactly.
RW_GROWTH = dict(h=40, nmlt=50, iring=500.0, beff=0.003, rho=1.7e-8, turns=4000, mode=1, amp=1e-6)


def write_rw_growth(d, g=RW_GROWTH):
    os.makedirs(d, exist_ok=True)
    text = cg.ring_conf(sections={
        "tracking": {"NrevTot": g["turns"], "NrevMon": 10, "EnableRW_long": 1, "Nmlt": g["nmlt"],
                     "EnableQuantum": 0},
        "ring": {"Lc": cg.small_ring(g["h"]), "Iring": g["iring"], "Cx": 0.0, "Cz": 0.0, "rhorw": g["rho"],
                 "beffL1": g["beff"], "beffL2": g["beff"], "beffH3": g["beff"], "beffV3": g["beff"]},
        "distribution": {"Np": 2, "xxCM_offset": 0.0, "zzCM_offset": 0.0}})
    return cg.write(os.path.join(d, "rw_growth.conf"), text)


def rw_growth_bunches(g=RW_GROWTH, beta=10.0):
    """Initial bunches [h][6][2] of the growth ring: every particle at the origin
    but for a transverse offset in the pattern of coupled-bunch mode mu in both
    planes, x_k - i beta x'_k = a exp(2 pi i mu k / h)."""
    import numpy as np
    out = np.zeros((g["h"], 6, 2))
    for k in range(g["h"]):
        ph = 2 * np.pi * g["mode"] * k / g["h"]
        for c in (2, 4):
            out[k, c] = g["amp"] * np.cos(ph)
            out[k, c + 1] = -g["amp"] / beta * np.sin(ph)
    return out


def rw_growth_analytic(E0, Lc, Iring, Q, beff, rho, beta=10.0):
    """Growth rate [1/s] of the bunch action (twice the amplitude's) of the most
    unstable transverse resistive-wall coupled-bunch mode: the closed form of the
    reference's test, src/test_rw.c:10-27 (tau_to_minus1) doubled as at
    :78-79, with the total current Iring [mA], E0 [GeV], Lc [m], the pipe radius
    beff [m], the resistivity rho [Ohm m] and the fractional tune of Q."""
    import numpy as np
    w0 = 2.0 * np.pi * 2.997925e8 / Lc
    eps0 = 8.854187817e-12
    dq = np.fmod(Q, 1.0)
    return 2.0 * (beta * w0 * Iring * 1e-3 / (4.0 * np.pi * E0 * 1e9)) * (Lc / (2.0 * np.pi) / beff ** 3) * \
        np.sqrt(2.0 * rho / ((1.0 - dq) * w0 * eps0))


def mode_growth(stats, mode, first, beta=10.0):
    """Slope per turn of log |a_mu|^2 over the turns from `first` on, per
    transverse plane, where a_mu is the discrete Fourier component mu over the
    bunches of x - i beta x' (centroids from stats [turns][bunches][12])."""
    import numpy as np
    out = []
    for p in (1, 2):
        A = stats[:, :, p] - 1j * beta * stats[:, :, 6 + p]
        F = np.fft.fft(A, axis=1)[:, mode]
        t = np.arange(len(stats))
        out.append(float(np.polyfit(t[first:], np.log(np.abs(F[first:]) ** 2), 1)[0]))
    return out


def write_c1_seed(d, seed, turns=1000, iring=400.0, beff=0.006):
    """The small ring driven unstable (confgen.wl_c1 at iring mA, pipe radius
    beff), quantum excitation on, with its three generator seeds moved by
    10 seed."""
    path = cg.wl_c1(d, turns=turns, iring=iring, beff=beff)
    text = open(path).read()
    for key, base in (("iseed_LON", 1001), ("iseed_HOR", 2002), ("iseed_VER", 3003)):
        text = text.replace("%s = %d" % (key, base), "%s = %d" % (key, base + 10 * seed))
    return cg.write(path, text)


def action_growth(ampinv, first=500, mon=10):
    """Slope per turn of log of the bunch-averaged centroid action (ampinv_cm)
    over the monitor turns from `first` on, HOR and VER."""
    import numpy as np
    turns = np.arange(len(ampinv))
    sel = (turns + 1) % mon == 0
    a = ampinv[sel][..., 3:6].mean(1)
    t = turns[sel]
    s = t >= first
    return [float(np.polyfit(t[s], np.log(a[s, p]), 1)[0]) for p in (1, 2)]
