"""Batching: tissues of different sheets, cell counts and grids advanced in
one context (every launch covers all of them) end bit-identical to each
tissue run alone, and each agrees with the oracle (MI355X only)."""
import numpy as np
import pytest

from epi_amd import configs
from parity import compare_cells, compare_fields

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10


def batch_specs():
    return [
        configs.config_tiny(7, n_free=160),
        configs.config_tiny(8, n_free=260, side=20, grid=20, nz=28),
        configs.config_tiny(9, n_free=120, side=16, grid=12, nz=20, n_layers=2),
        configs.config_tiny(10, n_free=300, side=24, grid=18, nz=30, dead_top=True, n_layers=3),
    ]


def _single(native, spec, draws, steps, ca_every, tuning=None):
    t = native.Tissue(spec["geometry"])
    t.load(spec)
    t.set_rng(draws)
    if tuning:
        t.set.set_tuning(tuning)
    t.step(native.Driver(0, 0, 0, ca_every, 0), steps)
    return t


@pytest.mark.parametrize("tuning", [None, {"b_zb": 2, "b_rows": 1, "atp_zb": 8, "atp_rows": 2, "ca_group": 4, "xcd": 1}])
def test_batch_equals_single_runs(oracle_mod, tuning):
    from epi_amd import native
    specs = batch_specs()
    steps, ca_every = 40, 20
    draws = [native.mt19937_64_uniform(100 + i, 300000) for i in range(len(specs))]
    print(f"batched: {len(specs)} tissues, {steps} steps, Ca every {ca_every}, tuning {tuning}", flush=True)
    ts = native.TissueSet([s["geometry"] for s in specs])
    for t, s in enumerate(specs):
        ts.load(t, s)
        ts.set_rng(t, draws[t])
    if tuning:
        ts.set_tuning(tuning)
    ts.step([native.Driver(0, 0, 0, ca_every, 0) for _ in specs], steps)
    for t, s in enumerate(specs):
        st = ts.status(t)[0]
        assert st.err == 0, (t, st.err)
        one = _single(native, s, draws[t], steps, ca_every)
        assert one.error() == 0, (t, one.error())
        a, b = ts.get_cells(t), one.get_cells()
        d = compare_cells(a, b)
        assert "n" not in d, (t, d)
        bad = {k: v for k, v in d.items() if v != 0}
        if bad:
            # name the differing entries and the side the oracle takes
            o = oracle_mod.OracleTissue(s["geometry"])
            o.load(s)
            o.set_rng(draws[t])
            dr = oracle_mod.Driver(0, 0, 0, ca_every, 0)
            for _ in range(steps):
                o.step(dr)
            r = o.get_cells()
            info = []
            if "gj" in bad:
                ptr = a["conn_ptr"]
                for e in np.nonzero(a["gj"] != b["gj"])[0][:6]:
                    c = int(np.searchsorted(ptr, e, side="right") - 1)
                    nb = int(a["conn"][e])
                    info.append((int(c), int(e - ptr[c]), int(a["state"][c]), nb, int(a["state"][nb]),
                                 float(a["gj"][e]), float(b["gj"][e]), float(r["gj"][e])))
            # the whole row of the first differing cell in each run, and
            # whether a second single run in a fresh context repeats it
            rows = {}
            if info:
                c0 = info[0][0]
                for name, cells in (("batched", a), ("single", b), ("oracle", r)):
                    p0, p1 = cells["conn_ptr"][c0], cells["conn_ptr"][c0 + 1]
                    rows[name] = (cells["conn"][p0:p1].tolist(), cells["gj"][p0:p1].tolist())
            again = _single(native, s, draws[t], steps, ca_every).get_cells()
            rows["single_again_equal_to_batched"] = {k: v for k, v in compare_cells(a, again).items() if v != 0}
            raise AssertionError((t, bad, "cell, slot, state, nb, nb state, batched, single, oracle", info, rows))
        fa, fb = ts.get_fields(t), one.get_fields()
        for k in fa:
            assert np.array_equal(fa[k], fb[k]), (t, k)
        assert st.rng_pos == one.rng_cursor()
        assert st.ca_solves == one.status().ca_solves


def test_batch_against_oracle(oracle_mod):
    from epi_amd import native
    specs = batch_specs()
    steps, ca_every = 30, 15
    draws = [native.mt19937_64_uniform(200 + i, 300000) for i in range(len(specs))]
    ts = native.TissueSet([s["geometry"] for s in specs])
    for t, s in enumerate(specs):
        ts.load(t, s)
        ts.set_rng(t, draws[t])
    ts.step([native.Driver(0, 0, 0, ca_every, 0) for _ in specs], steps)
    for t, s in enumerate(specs):
        o = oracle_mod.OracleTissue(s["geometry"])
        o.load(s)
        o.set_rng(draws[t])
        drv = oracle_mod.Driver(0, 0, 0, ca_every, 0)
        for _ in range(steps):
            o.step(drv)
        assert o.error() == 0
        d = compare_cells(ts.get_cells(t), o.get_cells())
        assert "n" not in d, (t, d)
        for k, v in d.items():
            if isinstance(v, int):
                assert v == 0, (t, k, v)
            else:
                assert v <= FP_TOL, (t, k, v)
        f = compare_fields(ts.get_fields(t), o.get_fields())
        for k, v in f.items():
            assert (v == 0) if k.startswith("cmap") else (v <= FP_TOL), (t, k, v)
        assert ts.status(t)[0].rng_pos == o.rng_cursor()


def test_launch_shapes_do_not_change_results():
    """Every launch shape the autotuner may pick gives bit-identical tissues."""
    from epi_amd import autotune, native
    specs = batch_specs()[:2]
    steps, ca_every = 30, 15
    draws = [native.mt19937_64_uniform(300 + i, 300000) for i in range(len(specs))]

    def run(tuning):
        ts = native.TissueSet([s["geometry"] for s in specs])
        for t, s in enumerate(specs):
            ts.load(t, s)
            ts.set_rng(t, draws[t])
        ts.set_tuning(tuning)
        ts.step([native.Driver(0, 0, 0, ca_every, 0) for _ in specs], steps)
        out = [(ts.get_cells(t), ts.get_fields(t)) for t in range(len(specs))]
        ts.close()
        return out

    ref = run(autotune.DEFAULT)
    shapes = [dict(autotune.DEFAULT, ca_group=g, xcd=x, atp_zb=zb, atp_rows=rows, b_zb=zb, b_rows=rows, ca_fuse=f,
                   map_lanes=ml, b_xcd=1 - x, sweep_wave=int(ml == 4), side_fork=x)
              for g, x, (zb, rows), f, ml in [(4, 1, (1, 4), 0, 4), (8, 0, (2, 2), 1, 1), (8, 1, (8, 1), 0, 4),
                                              (16, 1, (16, 4), 1, 4), (4, 0, (1, 1), 1, 16), (4, 0, (2, 4), 2, 4), (8, 1, (8, 4), 2, 1)]]
    for tu in shapes:
        print(f"shape {tu}", flush=True)
        got = run(tu)
        for t in range(len(specs)):
            (ca, fa), (cb, fb) = ref[t], got[t]
            d = compare_cells(ca, cb)
            assert "n" not in d and all(v == 0 for v in d.values()), (tu, t, d)
            for k in fa:
                assert np.array_equal(fa[k], fb[k]), (tu, t, k)


def test_step_graph_equals_one_by_one_launches():
    """A step's renewal and build replayed from captured graphs (STEP_GRAPH 1)
    gives the tissues that launching them one kernel at a time gives.  The run
    replays each graph many times, with force launches, Ca solves and status
    read-backs between the replays, and the cell counts change."""
    from epi_amd import native
    specs = batch_specs()
    steps, ca_every = 60, 20
    draws = [native.mt19937_64_uniform(500 + i, 300000) for i in range(len(specs))]

    def run(graph):
        ts = native.TissueSet([s["geometry"] for s in specs])
        for t, s in enumerate(specs):
            ts.load(t, s)
            ts.set_rng(t, draws[t])
        ts.set_tuning({"step_graph": graph})
        assert ts.get_tuning()["step_graph"] == graph
        n0 = [ts.status(t)[0].n for t in range(len(specs))]
        drv = [native.Driver(0, 0, 0, ca_every, 0) for _ in specs]
        for _ in range(3):
            ts.step(drv, steps // 3)
        out = [(n0[t], ts.status(t)[0], ts.get_cells(t), ts.get_fields(t)) for t in range(len(specs))]
        ts.close()
        return out

    eager = run(0)
    graph = run(1)
    for t in range(len(specs)):
        (_, sa, ca, fa), (n0, sb, cb, fb) = eager[t], graph[t]
        print(f"tissue {t}: n {n0} -> {sb.n}, solves {sb.ca_solves}", flush=True)
        assert sa.err == 0 and sb.err == 0
        assert (sa.n, sa.rng_pos, sa.ca_solves, sa.next_uid) == (sb.n, sb.rng_pos, sb.ca_solves, sb.next_uid)
        d = compare_cells(ca, cb)
        assert "n" not in d and all(v == 0 for v in d.values()), (t, d)
        for k in fa:
            assert np.array_equal(fa[k], fb[k]), (t, k)
    # divisions moved the cell counts (and with them the rounded launch sizes)
    assert any(s.n != n0 for n0, s, _, _ in graph)


def test_ca_solve_graph_replays_equal_eager_solves():
    """Whole Ca solves replayed from one captured graph (STEP_GRAPH 1): a
    forced solve every 5 steps for 40 steps, the tissues compared with the
    eager run after every 5 steps, so that each replay follows eager work
    and read-backs of its own"""
    from epi_amd import native
    specs = batch_specs()[:2]
    draws = [native.mt19937_64_uniform(700 + i, 300000) for i in range(len(specs))]

    def start(graph):
        ts = native.TissueSet([s["geometry"] for s in specs])
        for t, s in enumerate(specs):
            ts.load(t, s)
            ts.set_rng(t, draws[t])
        ts.set_tuning({"step_graph": graph})
        return ts, [native.Driver(0, 0, 0, 5, 0) for _ in specs]

    (ea, da), (gr, dg) = start(0), start(1)
    for chunk in range(8):
        ea.step(da, 5)
        gr.step(dg, 5)
        for t in range(len(specs)):
            sa, sb = ea.status(t)[0], gr.status(t)[0]
            assert sa.err == 0 and sb.err == 0
            assert (sa.n, sa.ca_solves, sa.rng_pos) == (sb.n, sb.ca_solves, sb.rng_pos), (chunk, t)
            d = compare_cells(ea.get_cells(t), gr.get_cells(t))
            assert "n" not in d and all(v == 0 for v in d.values()), (chunk, t, d)
            fa, fb = ea.get_fields(t), gr.get_fields(t)
            for k in fa:
                assert np.array_equal(fa[k], fb[k]), (chunk, t, k)
    print(f"solves per tissue: {[gr.status(t)[0].ca_solves for t in range(len(specs))]}", flush=True)
    ea.close()
    gr.close()
