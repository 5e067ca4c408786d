"""Runtime parameters on the GPU (MI355X): a tissue given parameters other
than epi5_MT's runs as the oracle does with the same ones, per tissue of a
batch, through the step driver (with its Ca gating, turnover and forced
cornification), the Ca solve and both windows.  Bar as in the parity tests:
integer, index and event state bit-exact, every FP64 field within 1e-10."""
import numpy as np
import pytest

from epi_amd import configs
from parity import compare_cells, compare_fields

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10

# every physics value changed by a few percent (alternating sign), the
# integer thresholds and counts changed, a short Ca averaging time (206
# sub-steps at DT_Ca 0.0097) and an early turnover so that initialize_sc and
# the num_sc solves run inside the window
_SCALE = ["DT_Cell", "R_max", "R_der", "COMPRESS_FACTOR", "THRESH_SP", "LJ_THRESH", "THRESH_DEAD", "ca2p_init",
          "ex_inert_init", "gj_init", "FAC_MAP", "agki_max", "fac", "Kpp", "dp", "kbc", "Hb", "Cout", "kg", "gamma",
          "mu0", "mu1", "kmu", "para_b", "para_bb", "para_k1", "k_flux", "beta_zero", "CA_OUT", "thgra", "delta_th",
          "thpri", "kpa", "delta_K", "Kpri", "delta_I", "para_k2", "H0", "wd", "ca2p_du", "STIM11", "Kaa", "Da",
          "AIR_STIM", "eps_m", "DER_DER_CONST", "K_TOTAL", "K_DESMOSOME_RATIO", "Kspring", "Kspring_d", "KBEND",
          "delta_R_coef", "para_ljp2", "Kspring_division", "stoch_div_time_ratio", "S2", "accel_div", "eps_kb",
          "alpha_b", "S0", "eps_ks", "accel_diff", "alpha_k", "eps_kk", "S1", "ubar", "delta_lipid", "delta_sig_r1",
          "lipid_rel", "lipid", "delta_L_coef", "ADHE_CONST", "eps_L", "unpair_dist_coef", "kb", "DUR_ALIVE",
          "DUR_DEAD", "DB"]


def perturbed(base, sign=1, pct=0.03):
    d = base.to_dict()
    kv = {k: d[k] * (1 + sign * pct * (1 if i % 2 == 0 else -1)) for i, k in enumerate(_SCALE)}
    kv.update(DT_Ca=0.0097, Ca_avg_time=2.0, SW_THRESH=12, div_max=12, NUM_SC_INIT=2, T_TURNOVER=0.5,
              iage_kitei=0.05, IP3_init=0.01, DISA_conn_num_thresh=12.0)
    return kv


def _check(gt, ot, tag):
    assert ot.error() == 0, (tag, "oracle error", ot.error())
    assert gt.error() == 0, (tag, "device error", gt.error())
    d = compare_cells(gt.get_cells(), ot.get_cells())
    assert "n" not in d, (tag, d)
    for k, v in d.items():
        assert (v == 0) if isinstance(v, int) else (v <= FP_TOL), (tag, k, v)
    f = compare_fields(gt.get_fields(), ot.get_fields())
    for k, v in f.items():
        assert (v == 0) if k.startswith("cmap") else (v <= FP_TOL), (tag, k, v)
    assert gt.rng_cursor() == ot.rng_cursor(), tag


def _pair(O, spec, kv, draws=600_000):
    from epi_amd import native
    g = spec["geometry"]
    o = O.OracleTissue(g)
    t = native.Tissue(g)
    o.set_params(kv)
    t.set_params(kv)
    o.load(spec)
    t.load(spec)
    rng = O.mt19937_64_uniform(spec["seed"] or 1, draws)
    o.set_rng(rng)
    t.set_rng(rng)
    return o, t


def test_parameters_read_back(oracle_mod):
    from epi_amd import native
    spec = configs.config_tiny(3, n_free=100, nz=24)
    t = native.Tissue(spec["geometry"])
    d = t.get_params().to_dict()
    ref = native.params_default().to_dict()
    for k in ("LX", "LY", "LZ", "NX", "NY", "NZ"):
        ref[k] = d[k]
    assert d == ref
    kv = perturbed(t.get_params())
    t.set_params(kv)
    got = t.get_params().to_dict()
    assert all(got[k] == v for k, v in kv.items())


def test_build_keys_are_refused():
    from epi_amd import native
    spec = configs.config_tiny(3, n_free=100, nz=24)
    t = native.Tissue(spec["geometry"])
    t.load(spec)
    for kv in ({"NX": spec["geometry"].nx + 1}, {"LZ": 99.0}, {"N3": 100}, {"AREA_GRID_ORIGINAL": 2.5},
               {"STOCHASTIC": 0}, {"DT_Ca": 0.0}, {"Ca_avg_time": 0.001}):
        with pytest.raises(native.EpiError) as e:
            t.set_params(kv)
        assert e.value.code == 7, kv
    # a window fixes the parameters while it is open
    t.set.mech_begin()
    with pytest.raises(native.EpiError):
        t.set_params({"Kpp": 0.31})
    t.set.mech_end()
    t.set_params({"Kpp": 0.31})
    assert t.get_params().Kpp == 0.31


def test_driver_with_changed_parameters(oracle_mod):
    """300 steps of the driver on a tiny tissue with every physics value
    changed: the gated and num_sc Ca solves, turnover (initialize_sc at step
    52), divisions and the renewal rates all follow the changed values"""
    from epi_amd import native
    O = oracle_mod
    spec = configs.config_tiny(9, n_free=300, nz=32)
    kv = perturbed(native.params_default())
    o, t = _pair(O, spec, kv)
    do, dg = O.Driver(0, 1, 0, 60, 0), native.Driver(0, 1, 0, 60, 0)
    for k in range(3):
        print(f"changed parameters: steps {100 * k}-{100 * k + 99}", flush=True)
        t.step(dg, 100)
        for _ in range(100):
            o.step(do)
        _check(t, o, f"chunk {k}")
    s = t.status()
    assert dg.force_cornif == 0 and s.ca_solves >= 5
    # the solves ran Ca_ITR = (int)(2.0 / 0.0097) = 206 sub-steps each
    assert s.grid_points_ca % 206 == 0


def test_batch_with_per_tissue_parameters(oracle_mod):
    """three tissues in one context with their own parameters (two Ca
    sub-step counts, so a solve of all three runs in two groups): each equals
    its own single-tissue run bit for bit and the oracle within the bar"""
    from epi_amd import native
    O = oracle_mod
    specs = [configs.config_tiny(s, n_free=250, nz=28) for s in (12, 13, 14)]
    base = native.params_default()
    kvs = [{}, perturbed(base, 1), dict(perturbed(base, -1), Ca_avg_time=3.0)]
    ts = native.TissueSet([s["geometry"] for s in specs])
    singles, oracles = [], []
    for i, (s, kv) in enumerate(zip(specs, kvs)):
        ts.set_params(i, kv)
        ts.load(i, s)
        rng = O.mt19937_64_uniform(s["seed"], 400_000)
        ts.set_rng(i, rng)
        o, t = _pair(O, s, kv, draws=400_000)
        singles.append(t)
        oracles.append(o)
    drv = [native.Driver(0, 0, 0, 40, 0) for _ in specs]
    print("per-tissue parameters: batch of 3, 120 steps", flush=True)
    ts.step(drv, 120)
    for i, (t, o) in enumerate(zip(singles, oracles)):
        t.step(native.Driver(0, 0, 0, 40, 0), 120)
        d = O.Driver(0, 0, 0, 40, 0)
        for _ in range(120):
            o.step(d)
        a, b = ts.get_cells(i), t.get_cells()
        for k in ("state", "x", "y", "z", "ca2p", "ip3", "agek", "ageb", "ca2p_avg", "ex_inert", "gj", "conn"):
            assert np.array_equal(a[k], b[k]), (i, k)
        _check(t, o, f"tissue {i}")
    assert [ts.status(i)[0].ca_solves for i in range(3)] == [t.status().ca_solves for t in singles]


def test_windows_with_changed_parameters(oracle_mod):
    """the reaction-diffusion and mechanics windows read the tissue's other
    parameters (DT_Ca, the reaction laws, LJ_THRESH, DT_Cell, ...) as well as
    their own coefficients: forward and reverse agree with the oracle's"""
    from epi_amd import native
    O = oracle_mod
    spec = configs.config_tiny(8, n_free=220, side=20, grid=20, nz=28)
    kv = perturbed(native.params_default())
    o, t = _pair(O, spec, kv, draws=200_000)
    d = O.Driver(0, 0, 0, 0, 0)
    for _ in range(3):
        o.step(d)
    t.step(native.Driver(0, 0, 0, 0, 0), 3)
    _check(t, o, "before the windows")
    zz = t.calc_zzmax()
    assert zz == o.calc_zzmax()
    st = t.get_cells()["state"]
    afm = np.isin(st, (configs.ALIVE, configs.FIX, configs.MUSUME))
    n_sub = 30
    prm = dict(du=0.0105, dp=0.095, gj0=0.97, kpa=np.where(st == configs.ALIVE, 4.4, 6.2))
    tgt = o.rd_forward(zz, n_sub, **prm)
    t.set.rd_begin(zz)
    t.set.rd_forward([prm], n_sub)
    u = t.set.rd_get_ca2p(0)
    assert np.max(np.abs(u[afm] - tgt[afm]) / np.abs(tgt[afm])) <= FP_TOL
    base = dict(du=0.01, dp=0.1, gj0=0.99, kpa=prm["kpa"] * 1.02)
    lo, go = o.rd_adjoint(zz, n_sub, tgt, **base)
    loss, grads = t.set.rd_adjoint([base], n_sub, 10, [tgt])
    t.set.rd_end()
    assert loss[0] == pytest.approx(lo, rel=1e-10)
    for k in ("du", "dp", "gj0"):
        assert grads[0][k] == pytest.approx(go[k], rel=1e-8), k
    # mechanics window from the same state
    mp = dict(eps=0.0102, k_adh=3.05, k_spring=25.5, k_bend=0.51)
    xo = o.mech_forward(20, **mp)
    t.set.mech_begin()
    t.set.mech_forward([mp], 20)
    xg = t.set.mech_get_positions(0)
    assert np.array_equal(xo, xg)
    lo, go = o.mech_adjoint(20, xo * 1.0001, **native.MECH_DEFAULT)
    loss, grads = t.set.mech_adjoint([native.MECH_DEFAULT], 20, 5, [xo * 1.0001])
    t.set.mech_end()
    assert loss[0] == pytest.approx(lo, rel=1e-12)
    for k in go:
        assert grads[0][k] == pytest.approx(go[k], rel=1e-8), k


def test_generated_tissue_runs_as_the_oracle(oracle_mod):
    """init_gen's tissue (epi_intg/calc/init_gen.cpp:48-106) on a 24 x 24 x
    40 box with R_der 1.3 and changed rates, loaded through cman_init's
    default start on both sides: 150 steps with a solve every 50"""
    from epi_amd import deck, native
    O = oracle_mod
    kv = dict(perturbed(native.params_default()), LX=24.0, LY=24.0, LZ=40.0, NX=24, NY=24, NZ=40, R_der=1.3,
              T_TURNOVER=6000.0)
    p = native.params_default().replace(**kv)
    cols, (nmx, nmy) = native.init_gen(p, nfix=8, der=2)
    n = cols["state"].size
    full = {k: np.zeros(n) for k in deck.DECK_COLUMNS}
    full.update(cols)
    full["pair"] = np.full(n, -1)
    full["fix_origin"] = np.where(cols["state"] == configs.FIX, cols["fix_origin"], -1)
    g = configs.Geometry(nmx=nmx, nmy=nmy, nx=24, ny=24, nz=40, lx=24.0, ly=24.0, lz=40.0, max_cells=n + 4000)
    spec = deck.spec_from_columns(full, g)
    spec["seed"] = 5
    o, t = _pair(O, spec, kv, draws=300_000)
    assert t.get_cells()["n"] == n
    do, dg = O.Driver(0, 0, 0, 50, 0), native.Driver(0, 0, 0, 50, 0)
    t.step(dg, 150)
    for _ in range(150):
        o.step(do)
    _check(t, o, "generated tissue")
    assert t.status().ca_solves >= 3


def test_odd_substep_count_solve_of_part_of_a_batch(oracle_mod):
    """Ca_avg_time / DT_Ca = 309 (odd) and a forced solve of one tissue of
    two: the buffer parities are shared by the context, so the solved
    tissue's buffers trade contents instead; both tissues equal their own
    single-tissue runs and the oracle"""
    from epi_amd import native
    O = oracle_mod
    specs = [configs.config_tiny(s, n_free=200, nz=24) for s in (21, 22)]
    kv = {"Ca_avg_time": 3.09}
    ts = native.TissueSet([s["geometry"] for s in specs])
    pairs = []
    for i, s in enumerate(specs):
        ts.set_params(i, kv)
        ts.load(i, s)
        ts.set_rng(i, O.mt19937_64_uniform(s["seed"], 200_000))
        pairs.append(_pair(O, s, kv, draws=200_000))
    assert ts.get_params(0).Ca_avg_time == 3.09 and int(3.09 / 0.01) == 309
    every = [25, 0]
    ts.step([native.Driver(0, 0, 0, e, 0) for e in every], 60)
    for i, ((o, t), e) in enumerate(zip(pairs, every)):
        t.step(native.Driver(0, 0, 0, e, 0), 60)
        d = O.Driver(0, 0, 0, e, 0)
        for _ in range(60):
            o.step(d)
        a, b = ts.get_cells(i), t.get_cells()
        for k in ("ca2p", "ca2p_next", "ip3", "ip3_next", "ca2p_avg", "ex_inert", "agek"):
            assert np.array_equal(a[k], b[k]), (i, k)
        fa, fb = ts.get_fields(i), t.get_fields()
        for k in ("atp", "atp_next", "b", "b_next"):
            assert np.array_equal(fa[k], fb[k]), (i, k)
        _check(t, o, f"tissue {i}")
    assert ts.status(0)[0].ca_solves == 2 and ts.status(1)[0].ca_solves == 0
