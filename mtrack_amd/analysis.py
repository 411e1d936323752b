"""Host-side analysis of tracked statistics.

growth_rates: the exponential growth (or damping) rate of the bunch-averaged
centroid action of each ring, the quantity a coupled-bunch instability is
judged by: a least-squares line through log <ampinv_cm> over the monitor turns
from `first` on, as the reference's own RW test does with its linear
regression on the logarithm of the amplitude (src/test_rw.c:66-90).
"""
import numpy as np


def growth_rates(ampinv, nrev_mon, first=0, ring_rows=None):
    """ampinv: [turns][bunches][6] as mbt_ring_track returns it (ampinv[3] then
    ampinv_cm[3] on monitor turns).  ring_rows: list of (first bunch, end bunch)
    per ring (an Ensemble's `first` boundaries); None = one ring.  Returns an
    array [rings][2] of the HOR and VER slopes per turn of log(mean over the
    ring's bunches of ampinv_cm)."""
    a = np.asarray(ampinv)
    turns = np.arange(a.shape[0])
    mon = ((turns + 1) % nrev_mon == 0) & (turns >= first)
    return growth_rates_monitor(a[mon], turns[mon], ring_rows)


def growth_rates_monitor(ampinv_mon, turns_mon, ring_rows=None):
    """The same fit on the monitor-turn rows only: ampinv_mon [rows][bunches][6]
    taken at turns turns_mon."""
    a = np.asarray(ampinv_mon)
    t = np.asarray(turns_mon, dtype=float)
    if ring_rows is None:
        ring_rows = [(0, a.shape[1])]
    out = np.full((len(ring_rows), 2), np.nan)
    if len(t) < 2:
        return out
    for k, (b0, b1) in enumerate(ring_rows):
        for j, p in enumerate((4, 5)):
            y = a[:, b0:b1, p].mean(axis=1)
            ok = y > 0
            if ok.sum() >= 2:
                out[k, j] = np.polyfit(t[ok], np.log(y[ok]), 1)[0]
    return out
