"""The multithreaded build of the CPU restatement (the CPU baseline of
bench.py) computes the same step as the serial parity checker: the same
neighbour sets (in schedule-dependent order, as the reference's lock-free
lists), the same maps, and positions and fields equal up to the rounding of
a different accumulation order (the reference's CAS sums race the same way)
and voxel owners up to the race between overlapping stamps."""
import numpy as np
import pytest

from epi_amd import configs


def _lists(c):
    p = c["conn_ptr"]
    return [sorted(c["conn"][p[i]:p[i + 1]].tolist()) for i in range(c["n"])]


def test_mt_build_matches_serial_up_to_order(oracle_mod):
    O = oracle_mod
    spec = configs.config_small(1)
    spec["cells"]["ageb"][:] = 0.0        # no cell is ready to divide: no draws
    out = []
    for mt in (False, True):
        t = O.OracleTissue(spec["geometry"], mt=mt)
        t.load(spec)
        d = O.Driver(0, 0, 0, 0, 0)
        for _ in range(3):
            t.step(d)
        t.check()
        out.append((t.get_cells(), t.get_fields()))
    (a, fa), (b, fb) = out
    assert a["n"] == b["n"]
    assert _lists(a) == _lists(b)
    for k in ("x", "y", "z"):
        assert np.allclose(a[k], b[k], rtol=0, atol=1e-11), k
    # overlapping stamps race: the same voxels are claimed, by any of the
    # cells whose spheres cover them (serially the highest index)
    assert np.array_equal(fa["cmap2"], fb["cmap2"])
    assert np.array_equal(fa["cmap1"] >= 0, fb["cmap1"] >= 0)
    same = fa["cmap1"] == fb["cmap1"]
    assert same.mean() > 0.9
    # a voxel whose owner differs takes another cornification source, and
    # diffusion carries that a little way: elsewhere B agrees to rounding
    db = np.abs(fa["b"] - fb["b"])
    assert np.median(db[same]) == 0.0 and db.max() < 0.05
    assert np.mean(db < 1e-12) > 0.9
    assert np.array_equal(a["state"], b["state"]) and np.array_equal(a["dermis"], b["dermis"])
