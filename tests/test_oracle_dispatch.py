"""How often, and with which law, cell_interaction applies each pair
(cell_interaction.cpp:529-685; Q2 and Q21 of the quirk list), pinned with
two- and three-cell systems whose displacement after one cell_interaction is
DT * sum(coef * (p_c - p_o)) over the applications.  The laws themselves are
pinned in test_oracle_laws; here a missed, doubled or mis-lawed application
moves the cells by a different multiple of d.

- DER-DER: the higher index applies CI_der_to_der; the lower index fails
  no_double_count (cell.h:293-295) and falls to else -> CI_other (ljmain),
  so the pair moves by DT (coef_DD + coef_LJ) d (:551-561).
- ALIVE-MUSUME: once, from the ALIVE side (:563-579; _MUSUME_interaction has
  no ALIVE branch, :612-636).
- FIX-MUSUME: once, from the FIX side (:581-603).
- DER-ALIVE: once, from the DER side with CI_other; ALIVE skips DER (:569-570).
- MUSUME-MEMB: the dermis gets CI_mu_to_memb, any other membrane neighbour
  CI_other (memb_exist<0>, cell.h:255-268).
- ALIVE-ALIVE: once (no_double_count on the ALIVE side).
"""
import numpy as np
import pytest

from helpers import sheet_spec
from oracle import LAWS

DT = 0.01


def build(O, free, nder=0, L=16.0):
    # a 20 x 20 sheet at the configs' pitch 0.8: membrane pairs stay in the
    # bounded branches of CI_memb_to_memb
    spec = sheet_spec(free, nmx=20, nmy=20, L=L)
    spec["nder"] = nder
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    return t, spec["nmemb"]


def moved(t):
    """positions before and the displacement of one cell_interaction"""
    c0 = t.get_cells()
    t.cell_interaction()
    t.check()
    c = t.get_cells()
    p0 = np.stack([c0["x"], c0["y"], c0["z"]], 1)
    return p0, np.stack([c["xn"], c["yn"], c["zn"]], 1) - p0


def test_der_der_is_applied_twice(oracle_mod):
    # three branches of CI_der_to_der: near (d + 0.56 < 2.8), contact, far
    for d in (2.0, 2.5, 3.1):
        t, m = build(oracle_mod, [dict(state=6, x=5.0, y=5.0, z=10.0), dict(state=6, x=5.0 + d, y=5.0, z=10.0)],
                     nder=2)
        p0, dp = moved(t)
        c_dd = t.coef(LAWS["DER_DER"], m + 1, m)
        c_lj = t.coef(LAWS["OTHER"], m, m + 1)
        assert c_lj != 0.0 and c_dd != 0.0
        want = DT * (c_dd + c_lj) * d            # cell m + 1 sits at +d
        assert dp[m + 1, 0] == pytest.approx(want, rel=1e-12), d
        assert dp[m, 0] == pytest.approx(-want, rel=1e-12), d
        assert np.all(dp[m:m + 2, 1:] == 0)


@pytest.mark.parametrize("lo,hi,law,side", [
    (0, 7, "ALAIRDE", 0),     # ALIVE (lower index) - MUSUME: ALIVE side only
    (7, 0, "ALAIRDE", 1),     # MUSUME (lower index) - ALIVE: still ALIVE side only
    (4, 7, "FIXMU", 0),       # FIX - MUSUME: FIX side only
    (7, 4, "FIXMU", 1),
    (0, 0, "ALAIRDE", 1),     # ALIVE - ALIVE: once, from the higher index
    (1, 0, "ALAIRDE", 1),     # DEAD - ALIVE: once
])
def test_single_application(oracle_mod, lo, hi, law, side):
    # d = 3.0: beyond contact (2.8), inside the 1.2 sigma adhesion range
    d = 3.0
    t, m = build(oracle_mod, [dict(state=lo, x=5.0, y=5.0, z=10.0, agek=5.0),
                              dict(state=hi, x=5.0 + d, y=5.0, z=10.0, agek=5.0)])
    p0, dp = moved(t)
    a, b = (m, m + 1) if side == 0 else (m + 1, m)
    coef = t.coef(LAWS[law], a, b)
    assert coef != 0.0
    want = DT * coef * d                      # displacement of the cell at +d
    assert dp[m + 1, 0] == pytest.approx(want, rel=1e-12)
    assert dp[m, 0] == pytest.approx(-want, rel=1e-12)


def test_der_alive_once_from_der_side_with_ljmain(oracle_mod):
    d = 2.6                                   # inside contact: ljmain != 0
    t, m = build(oracle_mod, [dict(state=6, x=5.0, y=5.0, z=10.0), dict(state=0, x=5.0 + d, y=5.0, z=10.0)],
                 nder=1)
    p0, dp = moved(t)
    coef = t.coef(LAWS["OTHER"], m, m + 1)
    assert coef != 0.0
    assert dp[m + 1, 0] == pytest.approx(DT * coef * d, rel=1e-12)
    assert dp[m, 0] == pytest.approx(-DT * coef * d, rel=1e-12)


def test_fix_musume_pair_is_not_adhesion(oracle_mod):
    # a FIX and its own MUSUME daughter: the FIX side skips its pair in the
    # list (fix->pair != conn, :588-595) and the pair spring applies once from
    # the higher index (_pair_interaction, :639-643)
    d = 2.0
    t, m = build(oracle_mod, [dict(state=4, x=5.0, y=5.0, z=10.0, pair=1, spr_nat_len=1.0),
                              dict(state=7, x=5.0 + d, y=5.0, z=10.0, pair=0, spr_nat_len=1.0)])
    p0, dp = moved(t)
    coef = t.coef(LAWS["PAIR"], m + 1, m)
    assert coef == pytest.approx(-5.0 * (d - 1.0) / d, rel=1e-14)   # Kspring_division (ell = 1)
    assert dp[m + 1, 0] == pytest.approx(DT * coef * d, rel=1e-12)
    assert dp[m, 0] == pytest.approx(-DT * coef * d, rel=1e-12)


def _memb_terms(t, c, m_end, dermis_law):
    """DT * sum over c's membrane neighbours of coef * (p_c - p_o), with the
    dermis under dermis_law and every other one under CI_other"""
    cc = t.get_cells()
    row = cc["conn"][cc["conn_ptr"][c]:cc["conn_ptr"][c + 1]]
    L = t.geo.lx
    out = np.zeros(3)
    n_memb = 0
    for o in row:
        if o >= m_end:
            continue
        n_memb += 1
        law = dermis_law if o == cc["dermis"][c] else "OTHER"
        k = t.coef(LAWS[law], c, o)
        d = np.array([cc["x"][c] - cc["x"][o], cc["y"][c] - cc["y"][o], cc["z"][c] - cc["z"][o]])
        d[:2] = np.where(d[:2] > 0.5 * L, d[:2] - L, np.where(d[:2] <= -0.5 * L, d[:2] + L, d[:2]))
        out += DT * k * d
    return out, n_memb, cc["dermis"][c]


@pytest.mark.parametrize("state,dermis_law,rdt", [(7, "MU_MEMB", 2), (7, "MU_MEMB", 0), (4, "FIX_MEMB", 0)])
def test_membrane_neighbours_dermis_spring_others_ljmain(oracle_mod, state, dermis_law, rdt):
    # a basal cell resting on the sheet touches several membrane particles;
    # only the nearest (its dermis) gets the spring law.  The membrane side
    # applies nothing to it (_MEMB_interaction visits membrane slots only).
    t, m = build(oracle_mod, [dict(state=state, x=8.1, y=7.9, z=3.0, rest_div_times=rdt)], L=16.0)
    p0, dp = moved(t)
    want, n_memb, derm = _memb_terms(t, m, m, dermis_law)
    assert n_memb >= 2 and 0 <= derm < m
    # dp is a difference of positions near 8: absolute rounding ~1e-15
    assert np.allclose(dp[m], want, rtol=1e-12, atol=1e-13)
    assert np.linalg.norm(want) > 1e-4
