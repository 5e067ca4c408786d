"""cell_state_renew (cell_state_renew.cpp:17-443) through the oracle:
replayed-draw accounting, division geometry, removal as swap-with-tail,
alias detection, and the per-state ageing laws at special points."""
import numpy as np
import pytest

from epi_amd import configs
from helpers import sheet_spec, pdiff

DT = 0.01
DELTA_L = 0.01 * 1.4


def ready_count(c):
    st, pair, derm, ageb, rdt = c["state"], c["pair"], c["dermis"], c["ageb"], c["rest_div_times"]
    other = np.arange(c["n"]) >= c["nmemb"] + c["nder"]
    fix = other & (st == 4) & (derm >= 0) & (pair < 0) & (ageb >= 6.0 * 0.75)
    mus = (other & (st == 7) & ~((derm < 0) & (pair < 0)) & (rdt > 0) & (pair < 0)
           & (ageb >= 6.0 * 0.75))
    return int(fix.sum() + mus.sum())


def test_draws_divisions_and_direction(oracle_mod):
    spec = configs.config_small(1)
    t = oracle_mod.OracleTissue(spec["geometry"])
    t.load(spec)
    t.set_rng(oracle_mod.mt19937_64_uniform(1, 2_000_000))
    L = spec["geometry"].lx
    ndiv_total = 0
    for step in range(120):
        t.cell_interaction()
        t.pos_copy()
        pre = t.get_cells()
        cur0 = t.rng_cursor()
        t.cell_state_renew()
        post = t.get_cells()
        assert t.error() == 0
        ndiv = post["n"] - pre["n"] + 0   # no removals in this tissue
        ndiv_total += ndiv
        # one draw per ready cell, two per direction attempt (one attempt here)
        assert t.rng_cursor() - cur0 == ready_count(pre) + 2 * ndiv
        for d in range(pre["n"], post["n"]):
            p = post["pair"][d]
            assert post["pair"][p] == d and post["uid"][d] == pre["uid"].max() + 1 + (d - pre["n"])
            assert post["state"][d] == 7 and post["spr_nat_len"][p] == pytest.approx(DELTA_L + DT * 0.14)
            v = np.array([pdiff(post["x"][p], post["x"][d], L), pdiff(post["y"][p], post["y"][d], L),
                          post["z"][p] - post["z"][d]]) / DELTA_L
            assert np.linalg.norm(v) == pytest.approx(1.0, abs=1e-9)
            der = pre["dermis"][p]
            u = np.array([pdiff(pre["x"][p], pre["x"][der], L), pdiff(pre["y"][p], pre["y"][der], L),
                          pre["z"][p] - pre["z"][der]])
            assert abs(np.dot(v, u / np.linalg.norm(u))) < 1e-9
            assert post["ageb"][p] == 0 and post["ageb"][d] == 0
        t.cell_pos_periodic_fix()
        t.connect_cell()
    assert ndiv_total > 0


def test_removal_is_swap_with_tail(oracle_mod):
    spec = configs.config_tiny(21, n_free=300, side=16, grid=16, n_layers=3, dead_top=True, nz=40)
    c = spec["cells"]
    dead = np.nonzero(c["state"] == 1)[0]
    c["agek"][dead] = 31.35              # removable when isolated enough
    t = oracle_mod.OracleTissue(spec["geometry"])
    t.load(spec)
    t.set_rng(oracle_mod.mt19937_64_uniform(3, 200000))
    t.cell_interaction()
    t.pos_copy()
    pre = t.get_cells()
    lens = np.diff(pre["conn_ptr"])
    queue = [i for i in range(pre["nmemb"], pre["n"])
             if pre["state"][i] in (1, 8) and pre["agek"][i] >= 31.3 and lens[i] <= 11]
    t.cell_state_renew()
    post = t.get_cells()
    assert t.error() == 0 and len(queue) > 0
    ndiv = post["uid"].max() - pre["uid"].max()
    assert post["n"] == pre["n"] + ndiv - len(queue)
    removed = set(pre["uid"][queue].tolist())
    assert set(post["uid"].tolist()) == (set(pre["uid"].tolist()) - removed) | set(
        range(pre["uid"].max() + 1, pre["uid"].max() + 1 + ndiv))
    # survivors keep their fields; slots of removed cells hold the tail cells
    idx = {u: i for i, u in enumerate(post["uid"])}
    for i in range(pre["n"]):
        if pre["uid"][i] in removed:
            continue
        j = idx[pre["uid"][i]]
        assert post["state"][j] == pre["state"][i] or pre["state"][i] in (0, 7)
        assert post["radius"][j] == pre["radius"][i]


def test_remove_alias_is_reported(oracle_mod):
    # two isolated removable DEAD cells as the last two slots: the second is
    # moved into the first's slot before its own turn (cellmanager.cpp:338-352)
    free = [dict(state=1, x=3.0, y=3.0, z=9.0, agek=31.5), dict(state=1, x=11.0, y=11.0, z=9.0, agek=31.5)]
    spec = sheet_spec(free, L=16.0, grid=(16, 16, 16))
    t = oracle_mod.OracleTissue(spec["geometry"])
    t.load(spec)
    t.cell_state_renew()
    assert t.error() == 5


def test_single_removal_of_the_last_cell_is_fine(oracle_mod):
    free = [dict(state=0, x=3.0, y=3.0, z=9.0, agek=1.0), dict(state=1, x=11.0, y=11.0, z=9.0, agek=31.5)]
    spec = sheet_spec(free, L=16.0, grid=(16, 16, 16))
    t = oracle_mod.OracleTissue(spec["geometry"])
    t.load(spec)
    t.cell_state_renew()
    c = t.get_cells()
    assert t.error() == 0 and c["n"] == spec["nmemb"] + 1


def test_ageing_laws_at_special_points(oracle_mod):
    # ALIVE below threshold, ca2p_avg <= ca2p_init: agek += DT * 0.05 * 0.02
    # DEAD not removable: agek += DT * 0.05 * 0.1
    # MUSUME (dermis, rest_div_times 0): ageb += DT * 0.12 * (0.1 + 5 * min0(avg - 0.122))
    free = [dict(state=0, x=3.0, y=3.0, z=9.0, agek=1.0, ca2p_avg=0.1),
            dict(state=1, x=11.0, y=11.0, z=9.0, agek=10.0),
            dict(state=7, x=8.0, y=8.0, z=3.4, ageb=1.0, ca2p_avg=0.2, rest_div_times=0),
            dict(state=0, x=3.0, y=11.0, z=9.0, agek=22.0)]
    spec = sheet_spec(free, L=16.0, grid=(16, 16, 16))
    t = oracle_mod.OracleTissue(spec["geometry"])
    t.load(spec)
    m = spec["nmemb"]
    t.cell_state_renew()
    c = t.get_cells()
    assert c["agek"][m] == pytest.approx(1.0 + DT * 0.05 * 0.02, rel=1e-15)
    assert c["agek"][m + 1] == pytest.approx(10.0 + DT * 0.05 * 0.1, rel=1e-15)
    assert c["ageb"][m + 2] == pytest.approx(1.0 + DT * 0.12 * (0.1 + 5 * (0.2 - 0.122)), rel=1e-14)
    assert c["state"][m + 3] == 1 and c["sw"] == 1       # cornified at agek >= 22
    # the lipid laws vanish when agek << THRESH_SP (tanh -> -1)
    assert c["in_fat"][m] == pytest.approx(0.0, abs=1e-12)


def test_musume_without_dermis_becomes_alive(oracle_mod):
    free = [dict(state=7, x=8.0, y=8.0, z=12.0, ageb=5.0, rest_div_times=3)]
    spec = sheet_spec(free, L=16.0, grid=(16, 16, 16))
    t = oracle_mod.OracleTissue(spec["geometry"])
    t.load(spec)
    t.set_rng(oracle_mod.mt19937_64_uniform(1, 10))
    t.cell_state_renew()
    c = t.get_cells()
    assert c["state"][spec["nmemb"]] == 0 and t.rng_cursor() == 0


def test_lipid_laws_at_special_points(oracle_mod):
    # k_lipid_release = 0.25 * 0.2 * (1 + tanh((avg - ubar) / 0.1)) * (1 + tanh((agek - 3) / 0.1))
    # k_lipid         = 0.25 * 0.6 * (1 + tanh((ubar - avg) / 0.1)) * (1 + tanh((agek - 3) / 0.1))
    # (cell_state_renew.cpp:102-119, ubar = 0.45, THRESH_SP = 3).  At agek = 3 and
    # avg = ubar both tanh are 0: release 0.05, lipid 0.15.  Five units past
    # either switch the tanh are +-1 in double: release 0.2 and lipid 0 above
    # ubar, release 0 and lipid 0.6 below.  in_fat = 0.4, ex_fat = 0.3:
    #   in' = in + DT (lipid (1 - in) - release in), ex' = ex + DT release in,
    #   agek' = agek + DT 0.05 (0.02 + 2 min0(avg - 0.122)) (:66-72, :347-358)
    cases = [  # agek, avg, release, lipid
        (3.0, 0.45, 0.05, 0.15),
        (13.0, 5.45, 0.2, 0.0),
        (13.0, -4.55, 0.0, 0.6),
    ]
    free = [dict(state=0, x=2.0 + 5.0 * i, y=8.0, z=9.0, agek=a, ca2p_avg=u, in_fat=0.4, ex_fat=0.3)
            for i, (a, u, _, _) in enumerate(cases)]
    spec = sheet_spec(free, L=16.0, grid=(16, 16, 16))
    t = oracle_mod.OracleTissue(spec["geometry"])
    t.load(spec)
    m = spec["nmemb"]
    t.cell_state_renew()
    t.check()
    c = t.get_cells()
    for i, (a, u, rel, lip) in enumerate(cases):
        assert c["in_fat"][m + i] == pytest.approx(0.4 + DT * (lip * 0.6 - rel * 0.4), rel=1e-14), i
        assert c["ex_fat"][m + i] == pytest.approx(0.3 + DT * rel * 0.4, rel=1e-14), i
        assert c["agek"][m + i] == pytest.approx(a + DT * 0.05 * (0.02 + 2 * max(u - 0.122, 0.0)), rel=1e-14), i
