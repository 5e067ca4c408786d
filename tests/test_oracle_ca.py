"""calc_ca2p (ca2p.cpp:14-402) through the oracle: special points of the
per-cell laws, antisymmetry of the gap-junction exchange, the sub-step
average and the ATP sources."""
import numpy as np
import pytest

from helpers import sheet_spec

DT = 0.01


def tissue(O, free, fields=0.1, L=16.0):
    spec = sheet_spec(free, nmx=4, nmy=4, L=L, grid=(16, 16, 16))
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    shape = (17, 17, 17)
    f = np.full(shape, fields)
    t.set_fields(atp=f, atp_next=f, b=f, b_next=f)
    t.setup_map_lat()
    return t, spec["nmemb"]


def test_dead_ip3_closed_form(oracle_mod):
    # DEAD cell with ALIVE neighbours of equal IP3: p' = p (1 - DT Kpp)
    free = [dict(state=1, x=8.0, y=8.0, z=9.0, ip3=0.5)]
    for dx, dy in ((2.5, 0), (-2.5, 0), (0, 2.5)):
        free.append(dict(state=0, x=8.0 + dx, y=8.0 + dy, z=9.0, ip3=0.5, agek=1.0))
    t, m = tissue(oracle_mod, free)
    t.ca_substep(12)
    c = t.get_cells()
    assert c["ip3"][m] == pytest.approx(0.5 * (1 - DT * 0.3), rel=1e-15)


def _pair_vs_isolated(O, u1, u2, p1, p2, agek=3.0):
    near = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=agek, ca2p=u1, ip3=p1),
            dict(state=0, x=10.5, y=8.0, z=9.0, agek=agek, ca2p=u2, ip3=p2)]
    far = [dict(near[0]), dict(near[1], x=2.0, y=2.0)]
    out = []
    for free in (near, far):
        t, m = tissue(O, free)
        t.ca_substep(12)
        out.append((t.get_cells(), m))
    return out


def test_gap_junction_exchange_is_antisymmetric(oracle_mod):
    (cp, m), (ci, _) = _pair_vs_isolated(oracle_mod, 0.10, 0.30, 0.20, 0.05)
    # agek = THRESH_SP: IAG = 0.5; gj = 0.99 on both directed edges
    ex_u = (cp["ca2p"][m:m + 2] - ci["ca2p"][m:m + 2]) / DT
    ex_p = (cp["ip3"][m:m + 2] - ci["ip3"][m:m + 2]) / DT
    assert ex_u[0] == pytest.approx(0.01 * 0.5 * 0.99 * (0.30 - 0.10), rel=1e-9)
    assert ex_u[0] == pytest.approx(-ex_u[1], rel=1e-9)
    assert ex_p[0] == pytest.approx(0.1 * 0.5 * 0.99 * (0.05 - 0.20), rel=1e-9)
    assert ex_p[0] == pytest.approx(-ex_p[1], rel=1e-9)


def test_no_exchange_without_iag(oracle_mod):
    # agek << THRESH_SP: IAG -> 0, the neighbours do not change each other
    (cp, m), (ci, _) = _pair_vs_isolated(oracle_mod, 0.10, 0.30, 0.20, 0.05, agek=-30.0)
    assert np.allclose(cp["ca2p"][m:m + 2], ci["ca2p"][m:m + 2], rtol=0, atol=1e-17)


def test_gap_junction_relaxation_at_wd(oracle_mod):
    # |u_o - u| = wd: tanh(0) = 0, so gj' = gj + DT (0.5 - gj)
    free = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=1.0, ca2p=0.25),
            dict(state=0, x=10.5, y=8.0, z=9.0, agek=1.0, ca2p=0.35)]
    t, m = tissue(oracle_mod, free)
    t.ca_substep(12)
    c = t.get_cells()
    k = np.nonzero(c["conn"][c["conn_ptr"][m]:c["conn_ptr"][m + 1]] == m + 1)[0][0]
    g = c["gj"][c["conn_ptr"][m] + k]
    assert g == pytest.approx(0.99 + DT * (0.5 - 0.99), abs=1e-12)


def test_average_is_the_sum_of_substeps(oracle_mod):
    free = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=2.0, ca2p=0.2),
            dict(state=7, x=10.5, y=8.0, z=9.0, ca2p=0.15),
            dict(state=1, x=8.0, y=10.5, z=9.0, ca2p=0.3)]
    t, m = tissue(oracle_mod, free)
    t.set_cell_values(ca2p_avg=np.zeros(t.get_cells()["n"]))
    acc = np.zeros(3)
    for _ in range(6):
        t.ca_substep(12)
        c = t.get_cells()
        acc = acc + c["ca2p"][m:m + 3]
        assert np.array_equal(c["ca2p_avg"][m:m + 2], acc[:2])
    assert c["ca2p_avg"][m + 2] == 0.0          # DEAD: no average


def test_solve_end_state(oracle_mod):
    free = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=2.0, ca2p=0.2),
            dict(state=1, x=8.0, y=10.5, z=9.0, ca2p=0.3, ca2p_avg=0.4)]
    t, m = tissue(oracle_mod, free)
    t.calc_ca2p(9.0)
    c = t.get_cells()
    assert c["ca2p"][m + 1] == 0 and c["ca2p_next"][m + 1] == 0 and c["ca2p_avg"][m + 1] == 0
    assert 0 < c["ca2p_avg"][m] < 2 and np.all(c["diff_u"] == 0)


def test_atp_air_source(oracle_mod):
    # a DEAD cell touching an AIR cell: its voxels gain DT * AIR_STIM
    free = [dict(state=1, x=8.0, y=8.0, z=9.0), dict(state=8, x=10.5, y=8.0, z=9.0)]
    t, m = tissue(oracle_mod, free, fields=0.0)
    shape = (17, 17, 17)
    c1 = np.full(shape, -1, np.int32)
    c1[8, 8, 9] = m
    t.set_maps(cmap1=c1, cmap2=np.zeros(shape, np.uint8))
    t.ca_substep(12)
    a = t.get_fields()["atp"]
    assert a[8, 8, 9] == pytest.approx(DT * 0.1, rel=1e-15)
    assert a[8, 8, 8] == 0.0
