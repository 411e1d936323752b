"""Growth-rate fit (mtrack_amd.analysis) on closed forms."""
import numpy as np

from mtrack_amd.analysis import growth_rates


def test_growth_rates_recover_exponentials():
    turns, nb = 600, 7
    t = np.arange(turns)
    am = np.zeros((turns, 2 * nb, 6))
    # ring 0: bunches 0..6 grow at 2e-3 (HOR) and damp at -1e-3 (VER)
    # ring 1: bunches 7..13 grow at 5e-3 in both planes
    for b in range(nb):
        am[:, b, 4] = (1 + 0.1 * b) * 1e-12 * np.exp(2e-3 * t)
        am[:, b, 5] = 3e-13 * np.exp(-1e-3 * t)
        am[:, nb + b, 4] = am[:, nb + b, 5] = 1e-14 * np.exp(5e-3 * t)
    r = growth_rates(am, 10, 100, [(0, nb), (nb, 2 * nb)])
    assert np.allclose(r, [[2e-3, -1e-3], [5e-3, 5e-3]], rtol=1e-9, atol=1e-12)


def test_growth_rates_single_ring_and_short_runs():
    am = np.ones((5, 3, 6))
    assert np.isnan(growth_rates(am, 10)).all()  # no two monitor turns
    am = np.ones((40, 3, 6))
    assert np.allclose(growth_rates(am, 10), 0.0)
