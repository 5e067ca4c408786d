"""The initial-state generator (init_gen, epi_intg/calc/init_gen.cpp:23-106)
on the host: the library's against the reference's own output and against
the oracle's restatement.  Host code only (no GPU call).

The shipped deck epi5_MT/input_2layer_nfix16.dat (the fixture
tests/golden/deck_2layer_nfix16.npz) is init_gen(16, 8) under epi5_MT's
parameters, written by CellManager::output with 15 fixed decimals: the
generator must reproduce it to the last printed digit."""
import numpy as np
import pytest

from epi_amd import configs
from epi_amd import native as N


def _fmt(a):
    return np.array(["%.15f" % v for v in a])


def test_sheet_size_is_define_h(deck_spec):
    """CalcParams' MEMB_NUM_X / MEMB_NUM_Y under the defaults are epi5_MT's
    NMX = 100 and NMY_tri = 114 (define.h:72-78), the shipped deck's sheet"""
    assert N.sheet_size() == (100, 114)
    assert deck_spec["nmemb"] == 100 * 114
    assert N.sheet_size(N.params_default().replace(USE_TRI_MEMB=0)) == (100, 100)


def test_defaults_reproduce_the_shipped_deck(deck_spec):
    cols, (nmx, nmy) = N.init_gen(nfix=16, der=8)
    c = deck_spec["cells"]
    assert cols["state"].size == c["state"].size == 11400 + 2312 + 16
    assert np.array_equal(cols["state"], c["state"])
    for k in ("x", "y", "z", "radius"):
        assert np.array_equal(_fmt(cols[k]), _fmt(c[k])), k
    fix = cols["state"] == configs.FIX
    assert np.array_equal(cols["fix_origin"][fix], np.arange(16))
    assert np.array_equal(cols["fix_origin"][fix], c["fix_origin"][fix])


@pytest.mark.parametrize("kv", [
    {},
    {"USE_TRI_MEMB": 0},
    {"R_der": 0.5},
    {"COMPRESS_FACTOR": 3.0, "LX": 30.0, "LY": 40.0},
    {"R_memb": 0.8, "R_max": 1.6},
])
def test_library_equals_oracle(oracle_mod, kv):
    p = N.params_default().replace(**kv)
    cols, sheet = N.init_gen(p, nfix=12, der=3)
    ocols, osheet = oracle_mod.init_gen(p, nfix=12, der=3)
    assert sheet == osheet
    for k in cols:
        assert np.array_equal(cols[k], ocols[k]), (kv, k)


def _min_memb_dist(cols, p, j, z):
    m = cols["state"] == configs.MEMB
    dx = cols["x"][j] - cols["x"][m]
    dy = cols["y"][j] - cols["y"][m]
    dx = np.where(dx > 0.5 * p.LX, dx - p.LX, np.where(dx <= -0.5 * p.LX, dx + p.LX, dx))
    dy = np.where(dy > 0.5 * p.LY, dy - p.LY, np.where(dy <= -0.5 * p.LY, dy + p.LY, dy))
    dz = z - cols["z"][m]
    return np.sqrt(dx * dx + dy * dy + dz * dz).min()


def test_relaxation_stops_at_first_contact():
    """with R_der 0.5 the ring starts 2.4 above the sheet and 1.1 (R_max +
    R_der) = 2.09: each FIX descends in steps of 0.001 and stops at the first
    height within 2.09 of a membrane cell"""
    p = N.params_default().replace(R_der=0.5)
    cols, _ = N.init_gen(p, nfix=8, der=2)
    thr = 1.1 * (p.R_max + p.R_der)
    start = p.R_max + p.R_memb + 2 * 2.0 * p.R_der + p.R_memb
    for j in np.nonzero(cols["state"] == configs.FIX)[0]:
        z = cols["z"][j]
        assert z < start - 0.2
        assert _min_memb_dist(cols, p, j, z) < thr
        assert _min_memb_dist(cols, p, j, z + 0.001) >= thr - 1e-12


def test_a_fix_that_cannot_reach_the_sheet_is_an_error():
    """a sparse sheet (pitch 4) and a contact distance of 0.33: the first
    FIX, about 1.0 from its nearest membrane cell across, sinks below
    Z_LEVEL - R_max, which the reference throws on"""
    p = N.params_default().replace(COMPRESS_FACTOR=0.5, R_der=0.1, R_max=0.2)
    with pytest.raises(N.EpiError) as e:
        N.init_gen(p, nfix=4, der=1)
    assert "zz too low" in str(e.value)


def test_driver_init_gen_mode_writes_the_shipped_deck(tmp_path, deck_spec):
    """`epi_proc --init-gen FILE` (epi_intg's "i" mode, main.cpp:103-112:
    init_gen_output(FILE, 16, 8)) writes CellManager::output's 15-decimal
    text; its state, radius and position columns are the shipped deck's
    characters.  Host code only: the mode makes no device call."""
    import os
    import subprocess
    drv = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "epi_amd", "host", "epi_proc")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    out = tmp_path / "init.dat"
    r = subprocess.run([drv, "--init-gen", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rows = [ln.split() for ln in out.read_text().splitlines()]
    c = deck_spec["cells"]
    assert len(rows) == c["state"].size
    cols = {"state": 1, "radius": 2, "x": 6, "y": 7, "z": 8}
    for k, j in cols.items():
        got = np.array([r_[j] for r_ in rows])
        want = np.array([str(int(v)) for v in c[k]]) if k == "state" else _fmt(c[k])
        assert np.array_equal(got, want), k
    # the rows read back through the deck loader as the same tissue
    from epi_amd import deck
    back = deck.spec_from_columns(deck.parse_deck(str(out)))
    assert back["nmemb"] == 11400 and back["nder"] == 2312
