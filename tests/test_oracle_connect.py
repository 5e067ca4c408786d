"""connect_cell (cell_connection.cpp:38-203): neighbour sets against brute
force, the canonical list order, gap-junction carry-over and the dermis."""
import numpy as np
import pytest

from epi_amd import configs
from helpers import lists, pdiff

AG = 2.0 + 1e-7


def tiny(O, seed=3, n_free=200, side=20):
    # L = 16 (a multiple of the bucket width): the 5x5x5 sweep covers every
    # contact across the periodic seam
    spec = configs.make_tissue(seed, nmx=side, nmy=side, lx=16.0, ly=16.0, nx=16, ny=16, nz=32,
                               n_fix=4, n_free=n_free, n_layers=2, lz=32.0)
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    return t, spec


def brute(c, L):
    n, nm = c["n"], c["nmemb"]
    x, y, z, r = c["x"], c["y"], c["z"], c["radius"]
    out = [set() for _ in range(n)]
    for i in range(nm, n):
        dx = pdiff(x[i], x[:i], L)
        dy = pdiff(y[i], y[:i], L)
        dz = z[i] - z[:i]
        rs = r[i] + r[:i]
        hit = np.nonzero(dx * dx + dy * dy + dz * dz <= 1.2 * 1.2 * rs * rs)[0]
        for j in hit:
            out[i].add(int(j))
            out[int(j)].add(i)
    return out


@pytest.mark.parametrize("seed,n_free", [(3, 200), (5, 400)])
def test_sets_match_brute_force(oracle_mod, seed, n_free):
    t, spec = tiny(oracle_mod, seed, n_free)
    c = t.get_cells()
    L = spec["geometry"].lx
    got = lists(c)
    bf = brute(c, L)
    nm = c["nmemb"]
    for i in range(c["n"]):
        extra = set(got[i][6:]) if i < nm else set(got[i])
        assert extra == bf[i], i
        assert len(got[i]) == len(set(got[i]))


def test_canonical_order(oracle_mod):
    t, spec = tiny(oracle_mod, 7, 300)
    c = t.get_cells()
    g = spec["geometry"]
    ANX, ANY, ANZ = g.buckets
    nm = c["nmemb"]
    x, y, z = c["x"], c["y"], c["z"]
    L = g.lx
    # grid_init bucket of every cell (cell_connection.cpp:44-46)
    bx = ((0.5 * L - pdiff(0.5 * L, x, L)) / AG).astype(int)
    by = ((0.5 * L - pdiff(0.5 * L, y, L)) / AG).astype(int)
    bz = (np.maximum(z, 0) / AG).astype(int)
    for i, row in enumerate(lists(c)):
        if i < nm:
            assert row[:6] == c["memb_adj"][6 * i:6 * i + 6].tolist()
            upper = row[6:]
            assert upper == sorted(upper) and all(o > i for o in upper)
            continue
        k = sum(1 for o in row if o < i)
        assert all(o < i for o in row[:k]) and all(o > i for o in row[k:])
        assert row[k:] == sorted(row[k:])
        # lower part: sweep order (j, k, l from -2..2 around the cell), then index
        ax, ay, az = int(x[i] / AG), int(y[i] / AG), int(z[i] / AG)
        def rank(o):
            dj = (bx[o] - ax + 2) % ANX
            dk = (by[o] - ay + 2) % ANY
            dl = (bz[o] - az + 2) % ANZ
            return (dj, dk, dl, o)
        assert row[:k] == sorted(row[:k], key=rank)


def test_gap_junctions_persist_and_reset(oracle_mod):
    t, spec = tiny(oracle_mod, 9, 200)
    c = t.get_cells()
    rng = np.random.default_rng(0)
    g = rng.uniform(0.1, 0.9, c["conn"].size)
    t.set_cell_values(gj=g)
    t.connect_cell()
    c1 = t.get_cells()
    assert np.array_equal(c1["conn"], c["conn"]) and np.array_equal(c1["gj"], g)
    # move one free cell away and back: its edges restart at gj_init = 0.99
    i = c["nmemb"] + 50
    x, y, z = c["x"].copy(), c["y"].copy(), c["z"].copy()
    z2 = z.copy()
    z2[i] = 14.0
    t.set_positions(x, y, z2)
    t.connect_cell()
    t.set_positions(x, y, z)
    t.connect_cell()
    c3 = t.get_cells()
    p = c3["conn_ptr"]
    old = {(a, int(b)): g[k] for a in range(c["n"]) for k, b in
           zip(range(c["conn_ptr"][a], c["conn_ptr"][a + 1]), c["conn"][c["conn_ptr"][a]:c["conn_ptr"][a + 1]])}
    for a in range(c3["n"]):
        for k in range(p[a], p[a + 1]):
            b = int(c3["conn"][k])
            if a == i or b == i:
                assert c3["gj"][k] == 0.99
            else:
                assert c3["gj"][k] == old[(a, b)]


def test_dermis_is_nearest_membrane_in_list(oracle_mod):
    t, spec = tiny(oracle_mod, 4, 300)
    c = t.get_cells()
    L = spec["geometry"].lx
    for i, row in enumerate(lists(c)):
        if c["state"][i] not in (4, 7):
            continue
        mem = [o for o in row if c["state"][o] == 9]
        if not mem:
            assert c["dermis"][i] == -1
            continue
        d = [pdiff(c["x"][i], c["x"][o], L) ** 2 + pdiff(c["y"][i], c["y"][o], L) ** 2
             + (c["z"][i] - c["z"][o]) ** 2 for o in mem]
        assert c["dermis"][i] == mem[int(np.argmin(d))]


def test_dense_sheet_has_long_rows(oracle_mod):
    """The dense-sheet tissue that the long-row force test runs has lists
    beyond the force kernel's 40 slots from its first build on."""
    from epi_amd import configs
    spec = configs.config_dense_sheet(11)
    ot = oracle_mod.OracleTissue(spec["geometry"])
    ot.load(spec)
    ot.set_rng(oracle_mod.mt19937_64_uniform(1, 100000))
    ot.connect_cell()
    ln = np.diff(ot.get_cells()["conn_ptr"])
    assert ot.error() == 0
    assert (ln > 40).sum() >= 10
