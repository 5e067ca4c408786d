"""cell_stat accounting on the GPU against the oracle (cellmanager.cpp:48-53,
338-352, 559-567; cell_state_renew.cpp:282-294): per-cell records and the
store of removed DEAD cells, exact (MI355X only)."""
import numpy as np
import pytest

from epi_amd import configs
from helpers import sheet_spec

pytestmark = pytest.mark.gpu


def test_two_isolated_cells():
    """an ALIVE cell past THRESH_DEAD and a MUSUME without dermis or
    partner, from driver step 100: records (-1, 100, 101) and
    (100, 101, 102), in removal order"""
    from epi_amd import native
    free = [dict(state=0, x=3.0, y=3.0, z=12.0, agek=31.5), dict(state=7, x=12.0, y=12.0, z=12.0, agek=40.0)]
    spec = sheet_spec(free, nmx=20, nmy=20, L=16.0, grid=(16, 16, 24), lz=24.0)
    t = native.Tissue(spec["geometry"])
    t.load(spec)
    t.set_rng(np.full(16, 0.5))
    m = spec["nmemb"]
    d = native.Driver(100, 0, 0, 0, 0)
    t.step(d, 1)
    cst = t.get_cells()["cst"]
    assert tuple(cst[m]) == (-1, 100, -1) and tuple(cst[m + 1]) == (100, -1, -1)
    t.step(d, 2)
    assert t.get_cells()["n"] == m and t.error() == 0
    assert t.take_cell_stats().tolist() == [[-1, 100, 101], [100, 101, 102]]
    assert t.take_cell_stats().shape == (0, 3)


def test_records_match_the_oracle_with_removals(oracle_mod):
    """a tissue whose top DEAD layer reaches the removal age over the run:
    every cell's record and the removed cells' store equal the oracle's"""
    from epi_amd import native
    spec = configs.config_tiny(21, n_free=300, side=16, grid=16, n_layers=3, dead_top=True, nz=40)
    c = spec["cells"]
    dead = np.nonzero(c["state"] == configs.DEAD)[0]
    c["agek"][dead] = np.linspace(31.28, 31.32, dead.size)
    alive = np.nonzero(c["state"] == configs.ALIVE)[0]
    hot = np.sort(alive[np.argsort(-c["z"][alive])[:40]])[:6]   # high, low indices: no remove alias
    c["agek"][hot] = np.linspace(31.35, 31.45, hot.size)    # cornify at once, removable next
    draws = oracle_mod.mt19937_64_uniform(21, 200000)
    o = oracle_mod.OracleTissue(spec["geometry"])
    o.load(spec)
    o.set_rng(draws)
    t = native.Tissue(spec["geometry"])
    t.load(spec)
    t.set_rng(draws)
    do, dg = oracle_mod.Driver(7, 0, 0, 0, 0), native.Driver(7, 0, 0, 0, 0)
    taken_o, taken_g = [], []
    for k in range(3):
        t.step(dg, 50)
        for _ in range(50):
            o.step(do)
        assert o.error() == 0 and t.error() == 0
        assert np.array_equal(t.get_cells()["cst"], o.get_cst()), k
        taken_g.append(t.take_cell_stats())
        taken_o.append(o.take_cell_stats())
        assert np.array_equal(taken_g[-1], taken_o[-1]), k
    allrec = np.concatenate(taken_o)
    assert allrec.shape[0] > 0 and np.any(allrec[:, 1] >= 7)   # removals of cells that cornified here
