"""The shipped input deck (epi5_MT/input_2layer_nfix16.dat) through the
oracle, against what the compiled reference was measured to do on it:
13,728 cells (11,400 MEMB, 2,312 DER, 16 FIX); after the first neighbour
build the mean list length is 6 + 0.98 for MEMB, 10.1 for DER, 36.7 for FIX
and 7.5 overall; the ext_stim box is 100 x 100 x 57; no division, removal or
Ca solve happens in the first steps."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def deck_tissue(oracle_mod, deck_spec):
    t = oracle_mod.OracleTissue(deck_spec["geometry"])
    t.load(deck_spec)
    return t


def test_deck_counts(deck_spec):
    st = deck_spec["cells"]["state"]
    assert st.size == 13728
    assert deck_spec["nmemb"] == 11400 and deck_spec["nder"] == 2312
    assert int(np.count_nonzero(st == 4)) == 16


def test_first_build_list_lengths(deck_tissue):
    c = deck_tissue.get_cells()
    ln = np.diff(c["conn_ptr"])
    st = c["state"]
    assert round(float(ln[st == 9].mean()) - 6, 2) == 0.98
    assert round(float(ln[st == 6].mean()), 1) == 10.1
    assert round(float(ln[st == 4].mean()), 1) == 36.7
    assert round(float(ln.mean()), 1) == 7.5
    # symmetric: every edge appears from both ends
    p, col = c["conn_ptr"], c["conn"]
    edges = {(i, int(j)) for i in range(c["n"]) for j in col[p[i]:p[i + 1]]}
    assert all((j, i) in edges for (i, j) in edges)


def test_ext_stim_box(deck_tissue):
    zz = deck_tissue.calc_zzmax()
    assert abs(zz - 25.8) < 1e-12
    assert int((zz + 2.0 * 1.4) * (200 / 100.0)) == 57


def test_first_steps_are_quiet(oracle_mod, deck_spec):
    t = oracle_mod.OracleTissue(deck_spec["geometry"])
    t.load(deck_spec)
    t.set_rng(oracle_mod.mt19937_64_uniform(1, 1000))
    d = oracle_mod.Driver(0, 1, 0, 0, 0)
    ran = sum(t.step(d) for _ in range(3))
    c = t.get_cells()
    assert t.error() == 0 and ran == 0
    assert c["n"] == 13728 and t.rng_cursor() == 0
