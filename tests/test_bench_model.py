"""bench.py's algorithmic byte model of the fused Ca launch, on a tissue's
lists from the oracle (no GPU): the per-cell and per-entry counts it uses are
recounted here by brute force from the cell list."""
import types

import numpy as np

import bench
from epi_amd import configs


def test_fused_byte_model_counts_coupled_entries(oracle_mod):
    spec = configs.config_tiny(7, n_free=200, nz=24)
    ot = oracle_mod.OracleTissue(spec["geometry"])
    ot.load(spec)
    ot.set_rng(oracle_mod.mt19937_64_uniform(1, 100000))
    ot.connect_cell()
    c = ot.get_cells()

    class Set:
        T = 1

        def get_cells(self, t):
            return c

    S = lambda solves, pts: types.SimpleNamespace(ca_solves=solves, grid_points_ca=pts, grid_points_b=0)
    launches, points = 999, 16 * 16 * 10
    r = bench.roofline(Set(), {"k_ca_fused": (10.0, 1000)}, [S(0, 0)], [S(1, points * 1000)], "k_ca_fused",
                       launches, "tiny")
    # brute force: AFM cells past the membrane and dermis, their coupled
    # neighbours (ALIVE, DEAD, FIX, MUSUME) and the ALIVE ones among them
    st, ptr, conn = c["state"], c["conn_ptr"], c["conn"]
    afm_states = (configs.ALIVE, configs.FIX, configs.MUSUME)
    coupled_states = (configs.ALIVE, configs.DEAD, configs.FIX, configs.MUSUME)
    cells = coupled = alive = 0
    for i in range(c["nmemb"] + c["nder"], st.size):
        if st[i] not in afm_states:
            continue
        cells += 1
        for o in conn[ptr[i]:ptr[i + 1]]:
            if st[o] in coupled_states:
                coupled += 1
                alive += st[o] == configs.ALIVE
    w = 999 / launches
    want = (bench.CA_CELL_BYTES * int(cells * w) + bench.CA_ENTRY_BYTES * coupled * w
            + bench.CA_ALIVE_ENTRY_BYTES * alive * w + bench.BYTES_PER_POINT["k_atp"] * points * w)
    assert cells > 0 and coupled > alive > 0
    assert r["algorithmic_bytes_per_launch"] == round(want, 1)
    assert r["avg_us"] == 10.0
    assert r["achieved"] == round(want / 10e-6 / 1e9, 2)
    assert r["units_per_launch"].startswith(f"{cells} AFM cells, {coupled / cells:.2f} coupled")
