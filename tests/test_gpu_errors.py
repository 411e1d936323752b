"""Error paths on the device side: requests the library refuses after it has
touched the device leave it usable (the next ring runs and gives the same
turns as a ring never preceded by the failure)."""
import numpy as np
import pytest

import mtrack_amd as m
from variants import write_variant

pytestmark = pytest.mark.gpu


def _turns(tmp_path, name="wakes"):
    d = str(tmp_path / ("ok_" + name))
    cfg = m.Config(write_variant(name, d), d + "/")
    g = m.Ring(cfg)
    g.prepare()
    st, _, _ = g.track(10)
    g.close()
    return st


def test_errors_leave_the_device_usable(tmp_path):
    before = _turns(tmp_path)
    # slices beyond the LDS histogram: refused after the allocations, which are freed
    d = str(tmp_path / "big")
    cfg = m.Config(write_variant("wakes", d, selffield={"Ncell": 3000}), d + "/")
    with pytest.raises(m.MbtError, match="too large for the LDS histogram"):
        m.Ring(cfg)
    # work decompositions outside the ring's range
    d = str(tmp_path / "chunk")
    cfg = m.Config(write_variant("wakes", d), d + "/")
    g = m.Ring(cfg)
    g.prepare()
    for items in (0, 128):
        with pytest.raises(m.MbtError, match="items per lane"):
            g.set_chunk(items)
    with pytest.raises(m.MbtError):
        g.get_bunch(g.Nb)
    st, _, _ = g.track(10)
    g.close()
    assert np.array_equal(st, before)
    assert np.array_equal(_turns(tmp_path), before)
