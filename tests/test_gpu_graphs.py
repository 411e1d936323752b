"""Turns replayed as HIP graphs against the same turns launched eagerly.

mbt_ring_track captures whole blocks of turns (a multiple of NrevMon, at most
256) as one graph and replays it; MBT_RING_NO_GRAPHS launches every kernel of
every turn.  Both must give bit-identical results, over several replays of one
graph, a partial block, and eager work (coordinate reads, potentials-style
reads) between the replays.
"""
import numpy as np
import pytest

import mtrack_amd as m
from variants import write_variant

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["wakes", "lr_fb", "scan", "qe"])
def test_graph_replays_equal_eager_turns(tmp_path, name):
    path = write_variant(name, str(tmp_path))
    print("case", name, flush=True)
    cfg = m.Config(path, str(tmp_path) + "/")
    a = m.Ring(cfg, graphs=True)
    b = m.Ring(cfg, graphs=False)
    for r in (a, b):
        r.prepare()
    for n in (250, 37, 500):
        sa, aa, ra = a.track(n)
        sb, ab, rb = b.track(n)
        assert np.array_equal(sa, sb) and np.array_equal(aa, ab) and np.array_equal(ra, rb), (name, n)
        assert np.array_equal(a.get_bunch(0), b.get_bunch(0))
        ha, ta = a.histograms()
        hb, tb = b.histograms()
        assert np.array_equal(ha, hb) and np.array_equal(ta, tb)
    assert np.array_equal(a.final_stats(), b.final_stats())
    a.close()
    b.close()


def test_timed_graphs(tmp_path):
    """Timing mode (per-kernel HIP events; launched eagerly, since the runtime
    does not time events recorded inside a graph) on a ring created with
    graphs: same results as an eager ring, and every kernel class of a turn
    with RW and LR resonators has a duration."""
    path = write_variant("lr_fb", str(tmp_path), tracking={"EnableRW_long": 1, "Nmlt": 2})
    cfg = m.Config(path, str(tmp_path) + "/")
    a = m.Ring(cfg, graphs=True)
    b = m.Ring(cfg, graphs=False)
    for r in (a, b):
        r.prepare()
        r.set_timing(True)
    sa, _, ra = a.track(40)
    sb, _, rb = b.track(40)
    assert np.array_equal(sa, sb) and np.array_equal(ra, rb)
    for r in (a, b):
        ms, n = r.kernel_times()
        print("kernel ms", ms, n)
        assert all(x == 40 for x in n)
        assert (ms > 0).all()
    a.close()
    b.close()


def test_graph_cache_keyed_by_monitor_offset(tmp_path):
    """A block of turns reuses a cached graph whenever its first monitor turn
    sits at the same offset, or when neither block holds a monitor turn: the
    pattern a benchmark uses (an untimed block, then timed blocks of the same
    length at other phases) instantiates no graph inside the timed blocks."""
    path = write_variant("wakes", str(tmp_path), tracking={"NrevMon": 100, "NrevTot": 200})
    cfg = m.Config(path, str(tmp_path) + "/")
    a = m.Ring(cfg, graphs=True)
    b = m.Ring(cfg, graphs=False)
    for r in (a, b):
        r.prepare()
    a.track(5)
    b.track(5)
    assert a.graph_builds() == 1
    a.track(20)  # turns 5..24: no monitor turn
    b.track(20)
    assert a.graph_builds() == 2
    sa, _, _ = a.track(20)  # turns 25..44: the same graph
    sb, _, _ = b.track(20)
    assert a.graph_builds() == 2
    assert np.array_equal(sa, sb)
    a.track(20)  # turns 45..64
    b.track(20)
    assert a.graph_builds() == 2
    sa, aa, _ = a.track(20)  # turns 65..84
    sb, ab, _ = b.track(20)
    assert a.graph_builds() == 2
    sa, aa, _ = a.track(20)  # turns 85..104: monitor turn 99 at offset 14, a new graph
    sb, ab, _ = b.track(20)
    assert a.graph_builds() == 3
    assert np.array_equal(sa, sb) and np.array_equal(aa, ab)
    assert np.any(aa[14] != 0) and not np.any(aa[13] != 0)
    a.close()
    b.close()


@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "eager"])
def test_async_blocks_equal_track(tmp_path, graphs):
    """mbt_ring_track_async + mbt_ring_track_wait, blocks of uneven length with
    potentials reads between them, give the bits mbt_ring_track gives; misuse is
    refused with MBT_EINVAL and leaves the ring usable."""
    path = write_variant("wakes", str(tmp_path))
    cfg = m.Config(path, str(tmp_path) + "/")
    a = m.Ring(cfg, graphs=graphs)
    b = m.Ring(cfg, graphs=graphs)
    for r in (a, b):
        r.prepare()
    amax = a.async_max()
    assert 1 <= amax <= 256
    for n in (9, 1, amax, 37):
        a.track_async(n)
        with pytest.raises(m.MbtError):
            a.track_async(1)  # one block at a time
        with pytest.raises(m.MbtError):
            a.track(1)  # tracking while a block is pending
        sa, aa, ra = a.track_wait()
        sb, ab, rb = b.track(n)
        assert np.array_equal(sa, sb) and np.array_equal(aa, ab) and np.array_equal(ra, rb), n
        pa, pb = a.potentials(0), b.potentials(0)
        assert all(np.array_equal(x, y) for x, y in zip(pa, pb))
    with pytest.raises(m.MbtError):
        a.track_wait()  # nothing pending
    with pytest.raises(m.MbtError):
        a.track_async(amax + 1)
    assert np.array_equal(a.get_bunch(0), b.get_bunch(0))
    assert np.array_equal(a.final_stats(), b.final_stats())
    a.close()
    b.close()
