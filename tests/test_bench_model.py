"""bench.py's algorithmic byte model of the fused Ca launch, on a tissue's
lists from the oracle (no GPU): the per-cell and per-entry counts it uses are
recounted here by brute force from the cell list."""
import types

import numpy as np
import pytest

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


def test_window_ends_on_a_forced_solve():
    # forced solves follow steps s with (s + 1) % ca_every == 0 (epi_step and
    # the oracle's eo_step); the timed steps are start + warmup .. + steps - 1
    for warmup, steps, every in ((5, 20, 500), (20, 5000, 500), (0, 1, 7), (3, 1000, 100), (10, 499, 500)):
        start = bench.window_start(warmup, steps, every)
        timed = range(start + warmup, start + warmup + steps)
        forced = [s for s in timed if (s + 1) % every == 0]
        assert forced and forced[-1] == timed[-1]
        assert len(forced) == -(-steps // every)
        # none in the untimed warm-up, unless the warm-up is a whole period long
        if warmup < every - (steps % every or every):
            assert not [s for s in range(start, start + warmup) if (s + 1) % every == 0]


def test_cpu_baseline_counts_each_solve_once(monkeypatch):
    # the per-step sample runs with sw gating held off, so the estimate is
    # steps * t_step + solves * 1000 * t_sub with nothing counted twice
    rates = {True: (0.010, 0.002), False: (0.040, 0.009)}
    monkeypatch.setattr(bench, "_cpu_rates", lambda O, spec, seed, mt, *a, **k: rates[mt])
    monkeypatch.setattr(bench, "cpu_threads", lambda: 16)
    spec = configs.config_tiny(1, n_free=40)
    r = bench.cpu_baseline(spec, 1, gpu_cell_steps=1000 * 20, gpu_ca_solves=1, timed_steps=20, budget_s=1.0)
    assert r["value"] == 1000 * 20 / (20 * 0.010 + 1 * 1000 * 0.002)
    assert r["cores"] == 16 and r["kind"] == "port"
    assert r["serial"]["value"] == 1000 * 20 / (20 * 0.040 + 1000 * 0.009)


def test_step_and_solve_byte_models(oracle_mod):
    spec = configs.config_tiny(7, n_free=200, nz=24)
    ot = oracle_mod.OracleTissue(spec["geometry"])
    ot.load(spec)
    c = ot.get_cells()
    n = c["n"]
    kbar = bench.mean_list_length(c)
    assert kbar == sum(c["conn_ptr"][i + 1] - c["conn_ptr"][i] for i in range(n)) / n
    assert bench.step_bytes(n, kbar, 1000) == n * (120 + 8 * kbar) + 26 * 1000
    # the split over kernel groups adds up to 120 + 8 k per cell
    assert sum(f(kbar) for _, _, _, f in bench.STEP_GROUPS) == pytest.approx(120 + 8 * kbar, rel=1e-14)
    # a solve: 1000 sub-steps of the per-sub-step model recounted by brute force
    st, ptr, conn = c["state"], c["conn_ptr"], c["conn"]
    cells = coupled = alive = 0
    for i in range(c["nmemb"] + c["nder"], n):
        if st[i] not in (configs.ALIVE, configs.FIX, configs.MUSUME):
            continue
        cells += 1
        for o in conn[ptr[i]:ptr[i + 1]]:
            if st[o] in (configs.ALIVE, configs.DEAD, configs.FIX, configs.MUSUME):
                coupled += 1
                alive += st[o] == configs.ALIVE
    want = 1000 * (96 * cells + 28 * coupled + 8 * alive + 22 * 500)
    assert bench.ca_solve_bytes(c, 500) == want


def test_dominant_kernel_is_the_costliest_modeled_one():
    prof = {"k_ca_fused": (1.0, 100), "k_sweep": (4.0, 10), "k_forces": (0.4, 10)}
    launched = {"k_ca_fused": 1000, "k_sweep": 20, "k_forces": 20}
    k, mix = bench.dominant_kernel(prof, launched)
    # estimated totals: fused 10 ms, sweep 8 ms (not modeled), forces 0.8 ms
    assert k == "k_ca_fused" and list(mix)[:2] == ["k_ca_fused", "k_sweep"]
    k, _ = bench.dominant_kernel({"k_sweep": (4.0, 10), "k_forces": (0.4, 10)}, launched)
    assert k == "k_forces"
