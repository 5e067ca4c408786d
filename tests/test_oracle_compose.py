"""The composition of one step (proc.cpp:22-32, 153-172) through the oracle's
eo_step: which call sees which other call's results.

- cell_interaction runs before cell_state_renew, so a daughter takes no
  force in the step that creates it: parent and daughter end the step
  exactly delta_L = 0.01 R_max apart, on the division axis
  (cell_state_renew.cpp:227-268).
- connect_cell runs after the renewal, so the daughter is in the same step's
  lists (both ends of the parent edge) and in its voxel map, where the
  highest index wins (map.cpp:91, 147).
- the periodic fix runs before connect_cell and setup_map_lat, so a cell
  that left the box is listed and stamped at its wrapped position.
- the cornification that brings sw to 20 triggers calc_ca2p in the same
  step (proc.cpp:165-172), and sw is reset after it; 19 do not.
"""
import numpy as np
import pytest

from epi_amd import configs
from helpers import sheet_spec

DELTA_L = 0.01 * 1.4


def tissue(O, free, L=16.0, grid=(16, 16, 16)):
    spec = sheet_spec(free, nmx=20, nmy=20, L=L, grid=grid)
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    return t, spec["nmemb"], spec


def test_daughter_takes_no_force_and_is_listed_and_stamped(oracle_mod):
    O = oracle_mod
    t, m, spec = tissue(O, [dict(state=4, x=8.1, y=7.9, z=2.6, ageb=10.0)])
    # one draw passes the division test, the next two give the direction
    t.set_rng(np.array([0.0, 0.3, 0.6, 0.5, 0.5, 0.5, 0.5]))
    n0 = t.get_cells()["n"]
    ran = t.step(O.Driver(0, 0, 0, 0, 0))
    t.check()
    c = t.get_cells()
    d = n0
    assert c["n"] == n0 + 1 and t.rng_cursor() == 3 and ran == 0
    assert c["state"][d] == configs.MUSUME and c["pair"][d] == m and c["pair"][m] == d
    sep = np.array([c["x"][m] - c["x"][d], c["y"][m] - c["y"][d], c["z"][m] - c["z"][d]])
    assert np.linalg.norm(sep) == pytest.approx(DELTA_L, rel=1e-9)
    # the axis is perpendicular to (parent - dermis)
    de = c["dermis"][m]
    nrm = np.array([c["x"][m] - c["x"][de], c["y"][m] - c["y"][de], c["z"][m] - c["z"][de]])
    assert abs(np.dot(sep, nrm)) / (np.linalg.norm(sep) * np.linalg.norm(nrm)) < 1e-2
    lists = [c["conn"][c["conn_ptr"][i]:c["conn_ptr"][i + 1]] for i in range(c["n"])]
    assert d in lists[m] and m in lists[d]
    assert lists[m][-1] == d                  # partners above the cell come last (ascending)
    f = t.get_fields()
    assert np.count_nonzero(f["cmap1"] == d) > 0
    # the daughter covers the parent's voxels around their common centre
    ix, iy, iz = c["lat"][3 * d:3 * d + 3]
    assert f["cmap1"][ix, iy, iz] == d


def test_wrapped_cell_is_listed_and_stamped_at_its_new_place(oracle_mod):
    # a cell just inside the x = L face, pushed out through it by a close
    # neighbour (ljmain at 1.5: a move of about +0.47 in this step)
    O = oracle_mod
    L = 16.0
    t, m, spec = tissue(O, [dict(state=0, x=L - 0.05, y=8.0, z=9.0, agek=1.0),
                            dict(state=0, x=L - 1.55, y=8.0, z=9.0, agek=1.0)])
    t.step(O.Driver(0, 0, 0, 0, 0))
    t.check()
    c = t.get_cells()
    assert 0.2 < c["x"][m] < 1.0                      # wrapped by the periodic fix
    assert tuple(c["lat"][3 * m:3 * m + 3]) == (0, 8, 9)
    f = t.get_fields()
    assert f["cmap1"][0, 8, 9] == m
    assert f["cmap1"][15, 8, 9] != m
    # the neighbour, now across the seam, is in its list
    row = c["conn"][c["conn_ptr"][m]:c["conn_ptr"][m + 1]]
    assert m + 1 in row


@pytest.mark.parametrize("ncorn,fires", [(20, True), (19, False)])
def test_twentieth_cornification_triggers_ca_in_the_same_step(oracle_mod, ncorn, fires):
    O = oracle_mod
    free = []
    for i in range(25):
        agek = 22.0 if i < ncorn else 5.0      # agek >= THRESH_DEAD cornifies this step
        free.append(dict(state=0, x=1.5 + 3.0 * (i % 5), y=1.5 + 3.0 * (i // 5), z=9.0, agek=agek,
                         ca2p=0.2, ca2p_avg=0.2))
    t, m, spec = tissue(O, free)
    t.set_fields(atp=np.full((17, 17, 17), 0.1), atp_next=np.full((17, 17, 17), 0.1),
                 b=np.full((17, 17, 17), 0.1), b_next=np.full((17, 17, 17), 0.1))
    ran = t.step(O.Driver(0, 0, 0, 0, 0))
    t.check()
    c = t.get_cells()
    dead = c["state"][m:m + 25] == configs.DEAD
    assert int(dead.sum()) == ncorn
    assert bool(ran) == fires
    if fires:
        assert c["sw"] == 0                                  # ca_calc_condition_reset
        assert np.all(c["ca2p"][m:m + ncorn] == 0) and np.all(c["ca2p_next"][m:m + ncorn] == 0)
        assert np.all(c["ca2p_avg"][m:m + ncorn] == 0)
        live = c["ca2p_avg"][m + ncorn:m + 25]
        assert np.all(live != 0.2) and np.all(live > 0)      # a solve's average
    else:
        assert c["sw"] == ncorn
        assert np.all(c["ca2p_avg"][m + ncorn:m + 25] == 0.2)
