"""cell_stat accounting in the oracle (cellmanager.cpp:48-53, 338-352,
559-567; cell_state_renew.cpp:282-294): a MUSUME with neither dermis nor
partner records k_aging_start when it turns ALIVE, an ALIVE cell records
k_cornified when it cornifies, and a DEAD cell's removal stores its record
with k_disap, in removal order; append_stat_data empties the store.

Two isolated cells far above the sheet, driver steps from 100:
- A, ALIVE with agek 31.5: cornifies at 100 (agek >= THRESH_DEAD), is
  removed at 101 (DEAD, agek >= 31.3, no neighbours): record (-1, 100, 101);
- B, MUSUME with agek 40: turns ALIVE at 100, cornifies at 101, is removed
  at 102: record (100, 101, 102).
A is removed first, and B (the tail) moves into its slot with its record."""
import numpy as np

from helpers import sheet_spec


def test_records_of_removed_dead_cells(oracle_mod):
    O = oracle_mod
    free = [dict(state=0, x=3.0, y=3.0, z=12.0, agek=31.5), dict(state=7, x=12.0, y=12.0, z=12.0, agek=40.0)]
    spec = sheet_spec(free, nmx=20, nmy=20, L=16.0, grid=(16, 16, 24), lz=24.0)
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    t.set_rng(np.full(16, 0.5))
    m = spec["nmemb"]
    a, b = m, m + 1
    assert np.all(t.get_cst() == -1)
    d = O.Driver(100, 0, 0, 0, 0)
    t.step(d)
    cst = t.get_cst()
    assert tuple(cst[a]) == (-1, 100, -1) and tuple(cst[b]) == (100, -1, -1)
    assert t.get_cells()["state"][a] == 1 and t.get_cells()["state"][b] == 0
    t.step(d)
    c = t.get_cells()
    assert c["n"] == m + 1                      # A removed, B moved into its slot
    assert tuple(t.get_cst()[a]) == (100, 101, -1)
    t.step(d)
    assert t.get_cells()["n"] == m
    rec = t.take_cell_stats()
    assert rec.tolist() == [[-1, 100, 101], [100, 101, 102]]
    assert t.take_cell_stats().shape == (0, 3)   # the store is emptied
    assert t.timestep() == 102 and t.error() == 0
    # membrane cells record nothing
    assert np.all(t.get_cst()[:m] == -1)
