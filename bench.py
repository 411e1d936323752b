#!/usr/bin/env python3
"""Benchmark: macroparticle-turns/s of the multi-bunch tracking turn on one MI355X.

Workload (production ring): 3 GeV, 354 m, h = 416, 416 bunches x 100,000
macroparticles, 100 longitudinal slices per bunch with integer counts,
broadband resonator (10 GHz, Q 1, 5 kOhm) short-range wake, resistive-wall
long-range wake over 416 buckets x 50 past turns, 0.5 mA/bunch, quantum
excitation off, FP64.  A step is one turn of the whole ring with every
coordinate resident in HBM; the initial distribution (the reference's
generator, made on the device) and set-up are outside the timed region.

Prints one JSON line.  `roofline` describes the per-particle pass kernel
k_turn: 96 algorithmic bytes per macroparticle-turn (read + write of six FP64
coordinates) over its average duration measured with HIP events on the ring's
stream; its `traffic` comes from a separate rocprofv3 PMC profile
(`traffic_source`).  The timed turns run the library's default path (each block
of turns replayed as one HIP graph, instantiated by an untimed block of the
same length first; `graphs_built_in_timed_region` counts instantiations inside
the timed call); the same number of turns is then launched eagerly with HIP
events around every kernel class (`instrumented`, `kernel_ms_per_turn`,
`roofline`), since the runtime does not time events inside a graph.
`cpu_baseline` times the reference's own turn loop (oracle/_ref/refharness,
built from the reference sources) on the same ring, the bunches of each turn
spread over the host's cores (`cores`), with a one-thread figure beside it;
without that binary, the project's C restatement on one core.

`batched` is the ensemble workload: 256 independent rings (seeded draws of
bunch count, macroparticle count, current and chromaticity; RF beam-loading
feedback on; every bunch from the reference generator) tracked together by
the same kernels, FP64, with its own throughput, k_turn roofline, placement
search and per-ring growth rates over further turns; `batched.mixed` repeats it
in mixed precision (FP32 wake convolutions, FP32 hi/lo MFMA resistive-wall
GEMM) with its errors against the FP64 run; `batched.one_ring_at_a_time_fp64`
tracks the same rings one after another through the single-ring FP64 path at
each ring's steady rate (the fixed cost of a call excluded), and
`vs_one_ring_at_a_time` is the batched forms' ratio to it, steady rate to
steady rate (`batched.steady`: a 60-turn call minus a 20-turn call, the same
outputs), with `vs_one_ring_at_a_time_per_call` the ratio of 20-turn calls,
fixed costs included on both sides.  `c1` is the
small ring (64 bunches x 1,000 macroparticles, QE on) in FP64 and mixed
precision (the lone ring's RW sum as the FP32 GEMM), with the mixed run's
errors and the growth rates of a resistive-wall instability in both.  `fbii`
is the weak-strong fast beam-ion workload (356-bunch train in h = 416), in
ion-slice interactions/s over one 100-turn call (the first turns after a host
synchronisation run slower, 4.0 falling to 3.4 ms in a kernel trace, so a
short call understates a long run's rate), with the ion-train kernel's FP64
roofline and the reference's own functions timed beside it.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_MP_TURN = 96.0
FP64_PEAK_TFLOPS = 78.6
# FP64 operations of one ion-slice interaction in the reference's train
# (src/fbi_weakstrong.c:370-441 with get_linearly_interpolated_Efields :454-580):
# two grid quotients and two fractions (8), six bilinear steps of a subtraction,
# a multiplication and an addition (18), the kicks of the ion (10) and the
# two field sums (4)
FLOPS_PER_INTERACTION = 40
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_turn_production_traffic.txt")
TRAFFIC_FILE_BATCHED = os.path.join(ROOT, "profiles", "k_turn_batched_traffic.txt")
TRAFFIC_NOTE = ("%s: HBM bytes per k_turn launch from separate rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of "
                "the same workload, scaled by the counters' reading on a plain device copy (not measured in this run)")


def production_conf(d, np_, turns):
    from mtrack_amd import confgen
    path = os.path.join(d, "production.conf")
    confgen.write(path, confgen.wl_production(turns=turns, np_=np_))
    return path


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per k_turn launch measured with rocprofv3 --pmc (scaled by the
    copy-kernel calibration), from the committed profile summary."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 2 and not line.startswith("#"):
                try:
                    vals[p[0]] = float(p[1])
                except ValueError:
                    pass
    return vals.get("traffic_bytes_per_launch")


def host_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS where set, else the
    machine's cores, at most 16 (the host share of one GPU)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)
    return max(1, min(n, 16))


def cpu_baseline(np_, turns=20, serial_turns=2, timeout=300):
    """The reference's own turn loop (oracle/_ref/refharness: its translation
    units compiled from the reference sources) on the production ring: the
    first `serial_turns` turns on one core, then `turns` turns with the bunches
    spread over the host's cores (one thread per bunch at a time, as the
    reference spreads them over MPI ranks).  Set-up is not timed.  Without the
    harness, the project's C restatement on one core."""
    harness = os.path.join(ROOT, "oracle", "_ref", "refharness")
    threads = host_threads()
    with tempfile.TemporaryDirectory() as d:
        conf = production_conf(d, np_, serial_turns + turns)
        if os.path.exists(harness):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            r = subprocess.run([harness, conf, d, str(serial_turns + turns), "0", "-1", str(serial_turns)],
                               capture_output=True, text=True, timeout=timeout, env=env)
            ser = par = None
            for line in r.stderr.splitlines():
                p = line.split()
                if p and p[0] == "serial_turns":
                    ser = float(p[3])
                elif p and p[0] == "parallel_turns":
                    par, used = float(p[3]), int(p[5])
            if r.returncode == 0 and ser and par:
                mp = 416 * np_
                return {"value": mp * turns / par, "unit": "macroparticle-turns/s", "cores": used,
                        "kind": "reference",
                        "sample": "production ring, 416 bunches x %d mp x %d turns, bunches over %d threads "
                                  "(turn loop only)" % (np_, turns, used),
                        "one_core": {"value": mp * serial_turns / ser, "cores": 1,
                                     "sample": "the same ring, %d turns on one thread" % serial_turns}}
        import mtrack_amd as m
        import oracle
        cfg = m.Config(conf, d + "/")
        orc = oracle.OracleRing(cfg)
        orc.prepare()
        t0 = time.perf_counter()
        orc.track(serial_turns)
        sec = time.perf_counter() - t0
        return {"value": 416 * np_ * serial_turns / sec, "unit": "macroparticle-turns/s", "cores": 1,
                "kind": "port", "sample": "production ring, 416 bunches x %d mp x %d turns" % (np_, serial_turns)}


def stat_errors(g, r):
    """Largest error of per-bunch statistics [..., 12]: centroids relative to
    the rms of the same coordinate, rms relative to itself."""
    import numpy as np
    g, r = np.asarray(g), np.asarray(r)
    e = 0.0
    for p in range(3):
        for mc, sc in ((p, 3 + p), (6 + p, 9 + p)):
            scale = np.maximum(np.abs(r[..., sc]), 1e-300)
            e = max(e, float(np.max(np.abs(g[..., mc] - r[..., mc]) / scale)),
                    float(np.max(np.abs(g[..., sc] - r[..., sc]) / scale)))
    return e


def growth_rates(am, nrev_mon, first):
    """Slope per turn of log(bunch-averaged centroid action), HOR and VER, over
    the monitor turns from `first` on."""
    import numpy as np
    turns = np.arange(len(am))
    mon = (turns + 1) % nrev_mon == 0
    a = am[mon][..., 3:6].mean(1)
    t = turns[mon]
    sel = t >= first
    return [float(np.polyfit(t[sel], np.log(a[sel, p]), 1)[0]) for p in (1, 2)]


def split_errors(g, r):
    """Errors of per-bunch statistics [..., 12] against a reference run: centroids
    (position and slope) relative to the reference rms of the same coordinate,
    rms relative to itself, and the emittance proxy rms(pos) * rms(slope)
    relative to itself, each the largest over bunches, turns and planes."""
    import numpy as np
    g, r = np.asarray(g), np.asarray(r)
    cen = rms = emit = 0.0
    for p in range(3):
        for mc, sc in ((p, 3 + p), (6 + p, 9 + p)):
            scale = np.maximum(np.abs(r[..., sc]), 1e-300)
            cen = max(cen, float(np.max(np.abs(g[..., mc] - r[..., mc]) / scale)))
            rms = max(rms, float(np.max(np.abs(g[..., sc] - r[..., sc]) / scale)))
        eg, er = g[..., 3 + p] * g[..., 9 + p], r[..., 3 + p] * r[..., 9 + p]
        emit = max(emit, float(np.max(np.abs(eg - er) / np.maximum(np.abs(er), 1e-300))))
    return {"centroid": cen, "rms": rms, "emittance": emit}


def c1(turns=2000):
    """The small multi-bunch ring (64 bunches x 1,000 macroparticles, QE on,
    copper RW, broadband resonator; confgen.wl_c1): the plain FP64 GPU path and
    the mixed-precision one (FP32 hi/lo-compensated wake convolutions and the
    long-range resistive-wall sum as an FP32 hi/lo MFMA GEMM, MBT_RING_RW_GEMM),
    each timed over `turns` turns from the same start, with the mixed run's
    centroid, rms and emittance errors against FP64; growth rates of the same
    ring driven unstable (400 mA, 6 mm RW radius, 1,000 turns) in both
    precisions."""
    import mtrack_amd as m
    from mtrack_amd import confgen
    out = {}
    with tempfile.TemporaryDirectory() as d:
        cfg = m.Config(confgen.wl_c1(d, turns=turns), d + "/")
        n = 64 * cfg.geti("Np")
        st = {}
        for mixed in (False, True):
            g = m.Ring(cfg, mixed=mixed, force_gemm=mixed)
            g.prepare()
            g.track(10)  # untimed: the turn graphs
            t0 = time.perf_counter()
            st[mixed], _, _ = g.track(turns)
            sec = time.perf_counter() - t0
            g.close()
            out["mixed" if mixed else "fp64"] = {"value": n * turns / sec, "unit": "macroparticle-turns/s",
                                                 "ms_per_step": 1e3 * sec / turns}
        out["mixed"]["error_vs_fp64"] = split_errors(st[True], st[False])
        cfg = m.Config(confgen.wl_c1(d, turns=1000, iring=400.0, beff=0.006), d + "/")
        rates = {}
        for mixed in (False, True):
            g = m.Ring(cfg, mixed=mixed, force_gemm=mixed)
            g.prepare()
            _, am, _ = g.track(1000)
            g.close()
            rates[mixed] = growth_rates(am, 10, 500)
        out["growth_rate_per_turn"] = {"fp64": rates[False], "mixed": rates[True],
                                       "rel_error": [abs(a - b) / abs(b) for a, b in zip(rates[True], rates[False])],
                                       "config": "400 mA, 6 mm RW radius, 1,000 turns, fit over turns 500-999"}
    out["steps"] = turns
    out["config"] = {"workload": "64 bunches (buckets int(6.5 k) of h 416) x 1,000 macroparticles, QE on, "
                                 "copper RW 12 mm over 19 turns, 10 GHz broadband resonator in 50 slices, 208 mA",
                     "precision": "mixed: hi/lo-compensated FP32 wake convolutions and the long-range "
                                  "resistive-wall sum as an FP32 hi/lo MFMA GEMM; coordinates, phase advance and "
                                  "RF recurrence FP64"}
    return out


def one_ring_at_a_time(nrings, seed, steps, warmup):
    """The batched workload's rings tracked one after another, each alone through
    the plain FP64 single-ring path (its own default decomposition and graphs),
    at each ring's steady rate: the ring's `steps`-turn and 3 x `steps`-turn
    calls are timed (both graphs instantiated first, at the same phase of the
    monitor interval as the timed calls) and the difference, 2 x
    `steps` turns without the fixed cost of a call (synchronisation, output
    copies, the first turns after an idle host phase), is summed over the rings.
    A user tracking thousands of turns per ring pays that fixed cost once."""
    import mtrack_amd as m
    from mtrack_amd import confgen
    total_s, total_mp, short_s = 0.0, 0, 0.0
    with tempfile.TemporaryDirectory() as d:
        for p in confgen.wl_ensemble(d, nrings=nrings, seed=seed):
            cfg = m.Config(p, d + "/")
            g = m.Ring(cfg)
            g.prepare()
            if warmup > 0:
                g.track(warmup)
            # untimed: the graphs of both block lengths, then a filler call back
            # to the same phase of the monitor interval, so that the timed calls
            # replay graphs already built (turn graphs are cached by block length
            # and the offset of the block's monitor turn)
            mon_every = cfg.geti("NrevMon")
            for _ in range(2):
                g.track(steps, want_amp=False)
                g.track(3 * steps, want_amp=False)
                fill = (-4 * steps) % mon_every
                if fill:
                    g.track(fill, want_amp=False)
                b0 = g.graph_builds()
                t0 = time.perf_counter()
                g.track(steps, want_amp=False)
                t1 = time.perf_counter()
                g.track(3 * steps, want_amp=False)
                t2 = time.perf_counter()
                if g.graph_builds() == b0:
                    break
            total_s += (t2 - t1) - (t1 - t0)
            short_s += t1 - t0
            total_mp += g.Nb * g.Np
            g.close()
    return {"value": total_mp * 2 * steps / total_s, "unit": "macroparticle-turns/s", "dtype": "f64",
            "rings": nrings, "steps": 2 * steps, "ms_per_step": 1e3 * total_s / (2 * steps),
            "per_call_value": total_mp * steps / short_s,
            "note": "each ring alone through the single-ring FP64 path; per ring, the time of a %d-turn call "
                    "minus that of a %d-turn call, summed over the rings (per_call_value: the %d-turn calls "
                    "alone, fixed costs included)" % (3 * steps, steps, steps)}


def placement(achieved, ceiling):
    """k_turn's rate against a plain in-place update of the same coordinate
    tiles (mbt_ring_stream_ceiling): the ceiling the tiles' placement in HBM
    sets for this traffic (it differs between allocations)."""
    return {"GB/s": ceiling, "k_turn_frac": achieved / ceiling,
            "what": "plain in-place update of the same coordinate tiles, bit-preserving, 96 B per slot"}


def batched(steps, warmup, nrings, seed, chunk=0, mixed=False, growth_turns=0):
    """The batched ensemble: nrings independent rings (bunch and macroparticle
    counts log-uniform, currents and chromaticities uniform, RF beam-loading
    feedback on; mtrack_amd.confgen.wl_ensemble) tracked together, FP64 (or
    mixed precision).  Returns the bench object and the statistics of the last
    timed turns."""
    import mtrack_amd as m
    from mtrack_amd import confgen
    with tempfile.TemporaryDirectory() as d:
        paths = confgen.wl_ensemble(d, nrings=nrings, seed=seed)
        cfgs = [m.Config(p, d + "/") for p in paths]
        # every bunch from the reference's generator, seeded per ring and bucket
        # (made on the device, bit for bit the host generator's)
        c0 = time.perf_counter()
        ens = m.Ensemble(cfgs, mixed=mixed)
        create_s = time.perf_counter() - c0
        tried, chosen = ens.placement()
        search = {"candidates": tried, "chosen_GB/s": chosen}
        ens.prepare()
        R, G, N = ens.size()
        if chunk:
            ens.set_chunk(chunk)
            tune_ms = None
        else:
            chunk, tune_ms = ens.autotune()
        if warmup > 0:
            ens.track(warmup)
        ens.track(steps)  # untimed: instantiates the graph of a `steps`-turn block
        # timed region: the default path, the turns replayed as HIP graphs
        builds0 = ens.graph_builds()
        t0 = time.perf_counter()
        last_stats, _, _ = ens.track(steps)
        t1 = time.perf_counter()
        builds_timed = ens.graph_builds() - builds0
        # instrumented pass: the same number of turns launched eagerly with HIP
        # events around every kernel class (per-kernel times, roofline)
        ens.set_timing(True)
        e0 = time.perf_counter()
        ens.track(steps)
        e1 = time.perf_counter()
        ms, launches = ens.kernel_times()
        bytes_turn = ens.bytes_per_turn()
        ens.set_timing(False)
        ceiling = ens.stream_ceiling(5)
        # steady rate, taken as one_ring_at_a_time takes its rings': a
        # 3 x steps-turn call minus a steps-turn call, the same outputs (no
        # ampinv).  Turn graphs are cached by block length and the offset of the
        # block's monitor turn, so the untimed pair instantiates both and an
        # untimed filler call brings the timed pair back to the same phase of
        # the monitor interval (checked: no graph built while timed)
        tracked = warmup + 3 * steps
        mon_every = cfgs[0].geti("NrevMon")
        steady = None
        for _ in range(2):
            ens.track(steps, want_amp=False)
            ens.track(3 * steps, want_amp=False)
            fill = (-4 * steps) % mon_every
            if fill:
                ens.track(fill, want_amp=False)
            b0 = ens.graph_builds()
            s0 = time.perf_counter()
            ens.track(steps, want_amp=False)
            s1 = time.perf_counter()
            ens.track(3 * steps, want_amp=False)
            s2 = time.perf_counter()
            tracked += 8 * steps + fill
            if ens.graph_builds() == b0:
                steady = (s2 - s1) - (s1 - s0)
                break
        growth = None
        if growth_turns > 0:
            # per-ring growth rates of the centroid action over further turns,
            # tracked in monitor-interval blocks keeping the monitor rows only
            from mtrack_amd.analysis import growth_rates_monitor
            import numpy as np
            ens.set_timing(False)
            mon = cfgs[0].geti("NrevMon")
            rows, tmon = [], []
            start = tracked  # turns tracked so far
            for blk in range(growth_turns // mon):
                if blk % 10 == 0:
                    print("growth-rate turns: %d of %d" % (blk * mon, growth_turns), file=sys.stderr, flush=True)
                _, am, _ = ens.track(mon, want_stats=False, want_rf=False)
                r = (mon - 1 - start % mon) % mon  # the block's monitor turn
                rows.append(am[r].copy())
                tmon.append(start + r)
                start += mon
            rates = growth_rates_monitor(np.array(rows), np.array(tmon),
                                         [(int(ens.first[k]), int(ens.first[k + 1])) for k in range(R)])
            growth = {"turns": growth_turns, "rings": R, "fit": "slope per turn of log <ampinv_cm>, HOR and VER",
                      "unstable_rings": int(np.sum(np.nanmax(rates, axis=1) > 0)),
                      "rate_per_turn": {"min": float(np.nanmin(rates)), "median": float(np.nanmedian(rates)),
                                        "max": float(np.nanmax(rates))},
                      "per_ring": [[round(float(v), 8) for v in r] for r in rates]}
        ens.close()
    sec = t1 - t0
    k_turn_ms = ms[0] / max(launches[0], 1)
    achieved = bytes_turn / (k_turn_ms * 1e-3) / 1e9
    return {"value": N * steps / sec, "unit": "macroparticle-turns/s", "dtype": "mixed" if mixed else "f64",
            "steps": steps,
            "warmup": warmup, "ms_per_step": 1e3 * sec / steps, "graphs_built_in_timed_region": builds_timed,
            "config": {"workload": "batched ensemble of %d independent rings in one launch per kernel: bunches "
                                   "32-416 and macroparticles 1e4-1e5 log-uniform, 0.1-5 mA/bunch, chromaticity "
                                   "-1..3, 100 slices, broadband resonator + resistive wall 50 turns + main cavity "
                                   "with RF beam-loading feedback, QE off (seed %d)" % (R, seed),
                       "rings": R, "bunches": G, "macroparticles": N,
                       "chunk_items": chunk, "autotune_ms": tune_ms},
            "kernel_ms_per_turn": {"k_turn": k_turn_ms, "k_reduce": ms[1] / max(launches[1], 1),
                                   "k_lr": ms[2] / max(launches[2], 1), "k_rw": ms[3] / max(launches[3], 1)},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(TRAFFIC_FILE_BATCHED),
                         "traffic_source": TRAFFIC_NOTE % os.path.relpath(TRAFFIC_FILE_BATCHED, ROOT)},
            "instrumented": {"note": "same number of turns launched eagerly, HIP events around every kernel class",
                             "ms_per_step": 1e3 * (e1 - e0) / steps, "value": N * steps / (e1 - e0)},
            "steady": ({"note": "a %d-turn call minus a %d-turn call, statistics and RF outputs: the fixed cost of "
                                "a call excluded, as in one_ring_at_a_time_fp64" % (3 * steps, steps),
                        "ms_per_step": 1e3 * steady / (2 * steps), "value": N * 2 * steps / steady}
                       if steady else None),
            "placement_ceiling": dict(placement(achieved, ceiling), search=search), "create_s": create_s,
            **({"growth_rates": growth} if growth else {})}, last_stats


def fbii_conf(d, train=356, np_=10000, turns=2000):
    """Fast beam-ion weak-strong workload: h = 416, a train of `train` bunches
    in buckets 0..train-1 followed by the gap, 2 Nslc + 1 = 41 slice centroids
    per bunch, 41 x 25 = 1,025 macro-ions created per bunch passage on the
    5 x 5 local grid of every slice, Bassetti-Erskine fields through the
    Faddeeva function, ion cloud cleared at the head of the train, FP64."""
    from mtrack_amd import confgen
    fill = os.path.join(d, "fill_fbii.txt")
    confgen.filling_file(fill, list(range(train)), out=())
    text = confgen.ring_conf(sections={
        "tracking": {"BunchModel": "STRONG", "EnableFBII": 1, "NrevTot": turns, "NrevMon": 100,
                     "EnableRW_long": 0, "EnableQuantum": 0, "filling": fill},
        "distribution": {"Np": np_, "Nslc": 20}})
    path = os.path.join(d, "fbii.conf")
    confgen.write(path, text)
    return path


def ion_slice_updates(nb, per_bunch=1025, slices=41):
    """Ion-slice interactions of one turn's train: bunch k meets the k*1025 ions
    of the bunches ahead of it at each of its 41 slices."""
    return sum(k * per_bunch for k in range(nb)) * slices


def fbii(steps, warmup, cpu=True):
    import mtrack_amd as m
    with tempfile.TemporaryDirectory() as d:
        cfg = m.Config(fbii_conf(d), d + "/")
        w = m.WeakStrong(cfg, bunch_currents=True)
        nb = w.Nb
        if warmup > 0:
            w.track(warmup)
        t0 = time.perf_counter()
        w.track(steps)
        t1 = time.perf_counter()
        # instrumented pass: HIP events around the ion train of every turn
        w.set_timing(True)
        w.track(steps)
        train_ms, trains = w.train_time()
        w.close()
    sec = (t1 - t0) / steps
    ions_ms = train_ms / max(trains, 1)
    tflops = ion_slice_updates(nb) * FLOPS_PER_INTERACTION / (ions_ms * 1e-3) / 1e12
    out = {"value": ion_slice_updates(nb) / sec, "unit": "ion-slice interactions/s", "dtype": "f64",
           "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * sec,
           "config": {"workload": "fast beam-ion weak-strong: h 416, %d-bunch train + %d-bucket gap, 41 slice "
                                  "centroids and 1,025 macro-ions per bunch passage, Bassetti-Erskine fields, "
                                  "bunch currents of the filling" % (nb, 416 - nb),
                      "bunches": nb, "ion_slice_interactions_per_turn": ion_slice_updates(nb)},
           "kernel_ms_per_turn": {"k_ws_ions": ions_ms},
           "roofline": {"bound": "fp64-valu", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tflops / FP64_PEAK_TFLOPS, "traffic": None,
                        "note": "k_ws_ions: %d FP64 operations per ion-slice interaction (the reference's loop, "
                                "src/fbi_weakstrong.c:370-441,454-580), a fused multiply-add counted as two; "
                                "duration from HIP "
                                "events around the kernel on its stream" % FLOPS_PER_INTERACTION}}
    if cpu:
        # the reference's beam-ion turn (oracle/_ref/refwsharness), turn loop only,
        # on a 128-bunch train: its train is a sequential relay through the
        # manager (one core), and its cost per ion-slice interaction does not
        # depend on the train length (interactions grow as the square of it)
        harness = os.path.join(ROOT, "oracle", "_ref", "refwsharness")
        if os.path.exists(harness):
            with tempfile.TemporaryDirectory() as d:
                conf = fbii_conf(d, train=128, turns=1)
                r = subprocess.run([harness, conf, d, "1", "1"], capture_output=True, text=True, timeout=300)
                sec = [float(l.split()[1]) for l in r.stderr.splitlines() if l.startswith("turn_loop_seconds")]
                if r.returncode == 0 and sec:
                    out["cpu_baseline"] = {"value": ion_slice_updates(128) / sec[0],
                                           "unit": "ion-slice interactions/s", "cores": 1, "kind": "reference",
                                           "sample": "128-bunch train + 288-bucket gap, 1 turn, turn loop only "
                                                     "(the reference relays the ion cloud through one process)"}
    return out


def copy_calibration():
    """Plain device copy of a known size, for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = 1 << 27  # 128 Mi doubles = 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--np", type=int, default=100000, dest="np_")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-turns", type=int, default=20, help="turns of the multi-core CPU baseline")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel eagerly (no HIP graphs)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensemble line")
    ap.add_argument("--batched-rings", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--batched-seed", type=int, default=20240607)
    ap.add_argument("--batched-chunk", type=int, default=0,
                    help="work items per lane for the batched line (0: autotune)")
    ap.add_argument("--batched-mixed", action="store_true", help="with --no-production: the batched line in mixed precision")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision batched run")
    ap.add_argument("--batched-growth-turns", type=int, default=2000,
                    help="after the timed turns, track this many more and report per-ring growth rates")
    ap.add_argument("--no-one-at-a-time", action="store_true",
                    help="skip timing the batched rings one at a time through the single-ring path")
    ap.add_argument("--no-c1", action="store_true", help="skip the small-ring FP64 / mixed line")
    ap.add_argument("--no-production", action="store_true",
                    help="skip the production line (profiling of the other lines only)")
    ap.add_argument("--no-fbii", action="store_true", help="skip the fast beam-ion line")
    ap.add_argument("--fbii-steps", type=int, default=100)
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return 0
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU (bunch-sharded) path is not part of it. "
              "Run with --gpus 1.", file=sys.stderr)
        return 2
    if a.no_production:
        out = {}
        if not a.no_batched:
            out["batched"] = batched(a.batched_steps, 3, a.batched_rings, a.batched_seed, a.batched_chunk,
                                     a.batched_mixed, a.batched_growth_turns)[0]
        if not a.no_fbii:
            out["fbii"] = fbii(a.fbii_steps, 5, cpu=False)
        print(json.dumps(out))
        return 0
    import mtrack_amd as m
    with tempfile.TemporaryDirectory() as d:
        conf = production_conf(d, a.np_, 20000)
        cfg = m.Config(conf, d + "/")
        ring = m.Ring(cfg, graphs=not a.no_graphs)
        ring.prepare()
        chunk, tune_ms = ring.autotune()
        nb, npb = ring.Nb, ring.Np
        if a.warmup > 0:
            ring.track(a.warmup)
        ring.track(a.steps)  # untimed: instantiates the graph of a `steps`-turn block
        # timed region: the default path, the turns replayed as HIP graphs (each
        # mbt_ring_track call ends with a stream synchronisation)
        builds0 = ring.graph_builds()
        t0 = time.perf_counter()
        ring.track(a.steps)
        t1 = time.perf_counter()
        builds_timed = ring.graph_builds() - builds0
        # instrumented pass: the same number of turns launched eagerly with every
        # kernel class bracketed by HIP events on the ring's stream
        ring.set_timing(True)
        e0 = time.perf_counter()
        ring.track(a.steps)
        e1 = time.perf_counter()
        ms, launches = ring.kernel_times()
        bytes_turn = ring.bytes_per_turn()
        ring.set_timing(False)
        ceiling = ring.stream_ceiling(10)
        ring.close()
    sec = t1 - t0
    value = nb * npb * a.steps / sec
    k_turn_ms = ms[0] / max(launches[0], 1)
    achieved = bytes_turn / (k_turn_ms * 1e-3) / 1e9
    out = {
        "metric": "macroparticle-turns/s on 1 MI355X (FP64 vs mixed/batched) + % HBM peak",
        "value": value,
        "unit": "macroparticle-turns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * sec / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: reference Gaussian generator (glibc rand stream), seeded",
        "config": {"workload": "production ring FP64: 416 bunches x %d macroparticles, 100 slices, "
                               "broadband resonator + resistive wall 416 x 50 turns, QE off" % npb,
                   "bunches": nb, "macroparticles_per_bunch": npb, "turns_per_step": 1,
                   "chunk_items": chunk, "autotune_ms": tune_ms},
        "kernel_ms_per_turn": {"k_turn": k_turn_ms, "k_reduce": ms[1] / max(launches[1], 1),
                               "k_lr": ms[2] / max(launches[2], 1), "k_rw": ms[3] / max(launches[3], 1)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "traffic_source": TRAFFIC_NOTE % os.path.relpath(TRAFFIC_FILE, ROOT)},
        "timed_path": "HIP graph replay of the turns (the library's default)" if not a.no_graphs
                      else "eager launches (--no-graphs)",
        "graphs_built_in_timed_region": builds_timed,
        "instrumented": {"note": "same number of turns launched eagerly, HIP events around every kernel class",
                         "ms_per_step": 1e3 * (e1 - e0) / a.steps, "value": nb * npb * a.steps / (e1 - e0)},
        "placement_ceiling": placement(achieved, ceiling),
    }
    if not a.no_c1:
        try:
            out["c1"] = c1()
        except Exception as e:
            out["c1"] = {"value": None, "error": str(e)[:300]}
    if not a.no_batched:
        try:
            out["batched"], st64 = batched(a.batched_steps, 3, a.batched_rings, a.batched_seed, a.batched_chunk,
                                           growth_turns=a.batched_growth_turns)
            if not a.no_mixed:
                mix, st32 = batched(a.batched_steps, 3, a.batched_rings, a.batched_seed,
                                    out["batched"]["config"]["chunk_items"], mixed=True)
                mix["error_vs_fp64"] = split_errors(st32, st64)
                for k in ("config", "roofline"):
                    mix.pop(k, None)
                out["batched"]["mixed"] = mix
            if not a.no_one_at_a_time:
                single = one_ring_at_a_time(a.batched_rings, a.batched_seed, a.batched_steps, 3)
                out["batched"]["one_ring_at_a_time_fp64"] = single
                # like for like: steady rates (fixed cost of a call excluded, the
                # same outputs) on both sides; the one-call ratio beside it
                # compares the timed steps-turn call with the rings' own
                # steps-turn calls, fixed costs included on both sides
                for form in [out["batched"]] + ([out["batched"]["mixed"]] if "mixed" in out["batched"] else []):
                    st = form.get("steady")
                    form["vs_one_ring_at_a_time"] = (st["value"] if st else form["value"]) / single["value"]
                    form["vs_one_ring_at_a_time_per_call"] = form["value"] / single["per_call_value"]
        except Exception as e:  # reported beside the FP64 line, never replaces it
            out["batched"] = {"value": None, "error": str(e)[:300]}
    if not a.no_fbii:
        try:
            out["fbii"] = fbii(a.fbii_steps, 5, cpu=not a.no_cpu_baseline)
        except Exception as e:
            out["fbii"] = {"value": None, "error": str(e)[:300]}
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(a.np_, a.cpu_turns)
        except Exception as e:  # the GPU number stands on its own
            out["cpu_baseline"] = {"value": None, "error": str(e)[:200]}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
