"""The benchmark workloads at their full sizes on the GPU (MI355X only),
against the oracle on the same seeded inputs:

- the large single tissue (128x128 sheet, 20,000 free cells, 128x128x256
  grid, seed 2) for 1000 steps with a forced calc_ca2p every 500 and the
  reference's sw gating, checked every 100 steps;
- the shipped deck (epi5_MT/input_2layer_nfix16.dat, 13,728 cells) for 200
  steps with a forced solve every 100;
- the batch of 256 tissues (sheet side U[32,96], free cells U[500,8000],
  grid edge U[32,96], nz = 2 edge, seeds 3000+t) for 20 steps ending in a
  forced solve of every tissue: each tissue bit-identical to its own
  single-tissue context, and sampled tissues against the oracle;
- the reaction-diffusion adjoint window (5,000 frozen free cells, 64x64x128
  grid, 200 sub-steps, checkpoint every 20): gradient against the oracle's
  reverse pass (1e-8) and central differences (1e-6);
- the batched mechanics adjoint (64 tissues of 2,000-6,000 cells, 48^3
  grids, 100-step windows, checkpoint every 10): sampled tissues against
  the oracle (1e-8) and central differences (1e-6), every tissue's gradient
  independent of the rest of the batch.

Bar as in the other parity tests: lists, events, maps and the random-draw
cursor bit-exact; every FP64 field within 1e-10 relative."""
import numpy as np
import pytest

from epi_amd import configs
from parity import compare_cells, compare_fields

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10
AFM = (configs.ALIVE, configs.FIX, configs.MUSUME)


def _check(a_cells, a_fields, a_cursor, o, tag):
    assert o.error() == 0, (tag, "oracle error", o.error())
    d = compare_cells(a_cells, o.get_cells())
    assert "n" not in d, (tag, d)
    for k, v in d.items():
        assert (v == 0) if isinstance(v, int) else (v <= FP_TOL), (tag, k, v)
    f = compare_fields(a_fields, o.get_fields())
    for k, v in f.items():
        assert (v == 0) if k.startswith("cmap") else (v <= FP_TOL), (tag, k, v)
    assert a_cursor == o.rng_cursor(), tag


LARGE_CHUNK = 100


@pytest.fixture(scope="module")
def large_pair(oracle_mod):
    """The large tissue in the oracle and on the GPU, advanced chunk by chunk
    by the tests below (each test is the next 100 steps of one run)."""
    from epi_amd import native
    O = oracle_mod
    spec = configs.config_large(2)
    g = spec["geometry"]
    draws = O.mt19937_64_uniform(2, 6_000_000)
    o = O.OracleTissue(g)
    o.load(spec)
    o.set_rng(draws)
    t = native.Tissue(g)
    t.load(spec)
    t.set_rng(draws)
    state = {"o": o, "t": t, "do": O.Driver(0, 0, 0, 500, 0), "dg": native.Driver(0, 0, 0, 500, 0), "done": 0,
             "n0": spec["cells"]["x"].size}
    yield state
    t.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("chunk", range(10))
def test_large_config_1000_steps(large_pair, chunk):
    """1000 steps of the large tissue with a forced solve after steps 499 and
    999 (plus whatever the sw gating fires), in chunks of 100 steps, each
    checked against the oracle: the chunks are one run, in order."""
    s = large_pair
    assert s["done"] == chunk * LARGE_CHUNK, "the chunks run in order"
    print(f"large config: steps {s['done']}-{s['done'] + LARGE_CHUNK - 1}", flush=True)
    s["t"].step(s["dg"], LARGE_CHUNK)
    for _ in range(LARGE_CHUNK):
        s["o"].step(s["do"])
    s["done"] += LARGE_CHUNK
    assert s["t"].error() == 0
    _check(s["t"].get_cells(), s["t"].get_fields(), s["t"].rng_cursor(), s["o"], f"large step {s['done'] - 1}")
    st = s["t"].status()
    assert st.steps == s["done"] and st.ca_solves >= s["done"] // 500
    if s["done"] == 1000:
        assert st.ndiv > 0 or s["t"].get_cells()["n"] != s["n0"]


@pytest.mark.timeout(600)
def test_shipped_deck_200_steps(oracle_mod, deck_spec):
    """The reference's own input deck through cman_init's default start
    (_cman_value_init, zero fields) for 200 steps, a forced solve every 100."""
    from epi_amd import native
    O = oracle_mod
    g = deck_spec["geometry"]
    draws = O.mt19937_64_uniform(16, 400_000)
    o = O.OracleTissue(g)
    o.load(deck_spec)
    o.set_rng(draws)
    t = native.Tissue(g)
    t.load(deck_spec)
    t.set_rng(draws)
    do, dg = O.Driver(0, 0, 0, 100, 0), native.Driver(0, 0, 0, 100, 0)
    for k in range(2):
        print(f"shipped deck: steps {100 * k}-{100 * k + 99}", flush=True)
        t.step(dg, 100)
        for _ in range(100):
            o.step(do)
        _check(t.get_cells(), t.get_fields(), t.rng_cursor(), o, f"deck step {100 * k + 99}")
    assert t.status().ca_solves >= 2


BATCH_STEPS, BATCH_CA_EVERY = 20, 20
ORACLE_SAMPLE = (0, 100, 200, 255)


@pytest.fixture(scope="module")
def batch256():
    """The 256 tissues advanced together in one context for 20 steps, the
    last one followed by a forced solve of every tissue."""
    from epi_amd import native
    specs = configs.batched_configs(256)
    print(f"batch of {len(specs)} tissues, {sum(s['cells']['x'].size for s in specs)} cells", flush=True)
    draws = [native.mt19937_64_uniform(s["seed"], 40 * (s["cells"]["x"].size + 64)) for s in specs]
    ts = native.TissueSet([s["geometry"] for s in specs])
    for t, s in enumerate(specs):
        ts.load(t, s)
        ts.set_rng(t, draws[t])
    ts.step([native.Driver(0, 0, 0, BATCH_CA_EVERY, 0) for _ in specs], BATCH_STEPS)
    out = []
    for t in range(len(specs)):
        st = ts.status(t)[0]
        out.append((st, ts.get_cells(t), ts.get_fields(t)))
    ts.close()
    return specs, draws, out


@pytest.mark.timeout(600)
@pytest.mark.parametrize("part", range(8))
def test_batch256_equals_single_runs(batch256, part):
    """tissues 32 part .. 32 part + 31 of the batch, each against its own
    single-tissue context over the same steps"""
    from epi_amd import native
    specs, draws, got = batch256
    for t in range(32 * part, 32 * part + 32):
        s = specs[t]
        st, cells, fields = got[t]
        assert st.err == 0, (t, st.err)
        assert st.ca_solves >= 1, t
        one = native.Tissue(s["geometry"])
        one.load(s)
        one.set_rng(draws[t])
        one.step(native.Driver(0, 0, 0, BATCH_CA_EVERY, 0), BATCH_STEPS)
        assert one.error() == 0, (t, one.error())
        d = compare_cells(cells, one.get_cells())
        assert "n" not in d, (t, d)
        bad = {k: v for k, v in d.items() if v != 0}
        assert not bad, (t, bad)
        f1 = one.get_fields()
        for k in fields:
            assert np.array_equal(fields[k], f1[k]), (t, k)
        s1 = one.status()
        assert (st.rng_pos, st.ca_solves, st.n, st.next_uid) == (s1.rng_pos, s1.ca_solves, s1.n, s1.next_uid), t
        one.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("t", ORACLE_SAMPLE)
def test_batch256_against_oracle(batch256, oracle_mod, t):
    specs, draws, got = batch256
    s = specs[t]
    st, cells, fields = got[t]
    print(f"batch256 tissue {t} ({s['cells']['x'].size} cells) against the oracle", flush=True)
    o = oracle_mod.OracleTissue(s["geometry"])
    o.load(s)
    o.set_rng(draws[t])
    drv = oracle_mod.Driver(0, 0, 0, BATCH_CA_EVERY, 0)
    for _ in range(BATCH_STEPS):
        o.step(drv)
    _check(cells, fields, int(st.rng_pos), o, f"batch256 tissue {t}")


@pytest.mark.timeout(600)
def test_rd_adjoint_full_window(oracle_mod):
    """The reaction-diffusion adjoint workload at its size: 64x64 sheet + 16
    FIX + 5,000 free cells, grid 64x64x128, 200 sub-steps, checkpoint every
    20, loss against a target from parameters perturbed by 5%."""
    from epi_amd import native
    O = oracle_mod
    spec = configs.config_adjoint_rd(4)
    g = spec["geometry"]
    draws = O.mt19937_64_uniform(4, 400_000)
    o = O.OracleTissue(g)
    o.load(spec)
    o.set_rng(draws)
    t = native.Tissue(g)
    t.load(spec)
    t.set_rng(draws)
    do = O.Driver(0, 0, 0, 0, 0)
    for _ in range(2):
        o.step(do)
    t.step(native.Driver(0, 0, 0, 0, 0), 2)
    zz = t.calc_zzmax()
    assert zz == o.calc_zzmax()
    st = t.get_cells()["state"]
    n = st.size
    afm = np.isin(st, AFM)
    rng = np.random.default_rng(4)
    kpa = np.where(st == configs.ALIVE, 4.5, 6.0) * (1 + 0.05 * rng.uniform(-1, 1, n))
    n_sub, every = 200, 20
    pert = dict(du=0.01 * 1.05, dp=0.1 * 0.95, gj0=0.99 * 0.95, kpa=kpa * (1 + 0.05 * rng.choice([-1, 1], n)))
    tgt = o.rd_forward(zz, n_sub, **pert)
    base = dict(du=0.01, dp=0.1, gj0=0.99, kpa=kpa)
    print("rd adjoint full window: oracle reverse pass", flush=True)
    lo, go = o.rd_adjoint(zz, n_sub, tgt, **base)
    t.set.rd_begin(zz)
    t.set.rd_forward([pert], n_sub)
    ug = t.set.rd_get_ca2p(0)
    assert np.max(np.abs(ug[afm] - tgt[afm]) / np.abs(tgt[afm])) <= FP_TOL
    print("rd adjoint full window: GPU reverse pass", flush=True)
    loss, grads = t.set.rd_adjoint([base], n_sub, every, [tgt])
    gg = grads[0]
    assert loss[0] == pytest.approx(lo, rel=1e-10)
    for k in ("du", "dp", "gj0"):
        assert gg[k] == pytest.approx(go[k], rel=1e-8), k
    sel = np.abs(go["kpa"]) > 1e-3 * np.max(np.abs(go["kpa"]))
    assert np.max(np.abs(gg["kpa"][sel] - go["kpa"][sel]) / np.abs(go["kpa"][sel])) <= 1e-8
    assert np.all(gg["kpa"][~afm] == 0)

    def L(**kw):
        p = dict(base)
        p.update(kw)
        t.set.rd_forward([p], n_sub)
        u = t.set.rd_get_ca2p(0)
        return np.mean((u[afm] - tgt[afm]) ** 2)

    for name, b in (("du", 0.01), ("dp", 0.1), ("gj0", 0.99)):
        h = b * 1e-5
        fd = (L(**{name: b + h}) - L(**{name: b - h})) / (2 * h)
        assert abs(gg[name] - fd) <= 1e-6 * abs(fd), (name, gg[name], fd)
    big = np.argsort(-np.abs(gg["kpa"]))[:3]
    for i in big:
        h = kpa[i] * 1e-5
        kp, km = kpa.copy(), kpa.copy()
        kp[i] += h
        km[i] -= h
        fd = (L(kpa=kp) - L(kpa=km)) / (2 * h)
        assert abs(gg["kpa"][i] - fd) <= 1e-6 * abs(fd), (i, gg["kpa"][i], fd)
    t.set.rd_end()


MECH_SAMPLE = (0, 21, 42, 63)
MECH_BASE = dict(eps=0.01, k_adh=3.0, k_spring=25.0, k_bend=0.5)
MECH_TARGET = dict(eps=0.0105, k_adh=2.85, k_spring=26.25, k_bend=0.475)


def _mech_perturb(c, seed):
    """curvature on the sheet and stretched FIX anchors, so that every
    parameter moves the loss well above rounding"""
    st = c["state"]
    rng = np.random.default_rng(seed)
    x, y, z = c["x"].copy(), c["y"].copy(), c["z"].copy()
    z[st == configs.MEMB] += rng.uniform(-0.15, 0.15, int((st == configs.MEMB).sum()))
    z[st == configs.FIX] += 0.6
    return x, y, z


@pytest.mark.timeout(900)
def test_mech_adjoint_full_batch(oracle_mod):
    """The batched mechanics adjoint workload at its size: 64 tissues of
    2,000-6,000 cells (48^3 grids), 100-step windows, checkpoint every 10."""
    from epi_amd import native
    O = oracle_mod
    specs = configs.adjoint_mech_configs(64)
    T = len(specs)
    n_steps, every = 100, 10
    ts = native.TissueSet([s["geometry"] for s in specs])
    draws = [O.mt19937_64_uniform(s["seed"], 300_000) for s in specs]
    for i, s in enumerate(specs):
        ts.load(i, s)
        ts.set_rng(i, draws[i])
    ts.step([native.Driver(0, 0, 0, 0, 0) for _ in specs], 2)
    starts = []
    for i, s in enumerate(specs):
        x, y, z = _mech_perturb(ts.get_cells(i), s["seed"])
        ts.set_positions(i, x, y, z)
        starts.append((x, y, z))
    ts.mech_begin()
    ts.mech_forward([MECH_TARGET] * T, n_steps)
    tgts = [ts.mech_get_positions(i) for i in range(T)]
    # per-tissue parameters, so that a tissue reading another's would show
    params = [{k: v * (1 + 0.002 * ((i % 7) - 3)) for k, v in MECH_BASE.items()} for i in range(T)]
    print("mech adjoint full batch: GPU reverse pass", flush=True)
    loss, grads = ts.mech_adjoint(params, n_steps, every, tgts)
    for i in MECH_SAMPLE:
        s = specs[i]
        print(f"mech adjoint full batch: tissue {i} ({s['cells']['x'].size} cells) against the oracle", flush=True)
        o = O.OracleTissue(s["geometry"])
        o.load(s)
        o.set_rng(draws[i])
        d = O.Driver(0, 0, 0, 0, 0)
        for _ in range(2):
            o.step(d)
        o.set_positions(*starts[i])
        xo = o.mech_forward(n_steps, **MECH_TARGET)
        assert np.array_equal(xo, tgts[i]), i
        lo, go = o.mech_adjoint(n_steps, tgts[i], **params[i])
        assert loss[i] == pytest.approx(lo, rel=1e-12), i
        for k in go:
            assert grads[i][k] == pytest.approx(go[k], rel=1e-8), (i, k)

    # central differences of the GPU window, every tissue at once (each
    # tissue's loss depends on its own parameters only).  The loss is only
    # piecewise smooth: adhesion has kinks at contact and at 1.2 sigma
    def losses(name, scale):
        p = [dict(q, **{name: q[name] * scale}) for q in params]
        ts.mech_forward(p, n_steps)
        out = []
        for i in range(T):
            d = ts.mech_get_positions(i) - tgts[i]
            out.append(np.mean(np.sum(d * d, axis=0)))
        return np.array(out)

    # Each tissue's difference is taken at two steps.  Their spread measures
    # the quotient's own error: FP64 atomics make every forward differ in its
    # last bits (which dominates for a parameter the loss barely depends on,
    # such as k_spring), and a kink of the force dispatch inside the step
    # shows the same way.  Every tissue's quotient agrees with the adjoint
    # within 1e-5 plus three spreads, and where the spread is below 1e-7 (at
    # least 8 tissues needed) the median agreement is 1e-7.  A missing or
    # wrong term of the adjoint is off by far more (the sampled tissues match
    # the oracle's adjoint to 1e-8 above).
    for name in MECH_BASE:
        scale = np.array([q[name] for q in params])
        fd = {}
        for eps in (1e-6, 5e-7):
            lp, lm = losses(name, 1 + eps), losses(name, 1 - eps)
            fd[eps] = (lp - lm) / (2 * eps * scale)
        noise = np.abs(fd[1e-6] - fd[5e-7]) / np.abs(fd[1e-6])
        g = np.array([gr[name] for gr in grads])
        rel = np.abs(g - fd[1e-6]) / np.abs(fd[1e-6])
        bad = rel > 1e-5 + 3 * noise
        assert not bad.any(), (name, np.nonzero(bad)[0], rel[bad], noise[bad])
        quiet = noise <= 1e-7
        print(f"mech adjoint full batch: {name}: {np.count_nonzero(quiet)} quiet tissues, "
              f"median spread {np.median(noise):.2e}, max agreement {rel.max():.2e}", flush=True)
        if np.count_nonzero(quiet) >= 8:
            assert np.median(rel[quiet]) <= 1e-7, (name, float(np.median(rel[quiet])))
    # a tissue's gradient does not depend on the rest of the batch: the same
    # window with only tissue 0's parameters changed leaves the others' alone
    p2 = [dict(q) for q in params]
    p2[0] = dict(MECH_TARGET)
    _, grads2 = ts.mech_adjoint(p2, n_steps, every, tgts)
    for i in range(1, T):
        for k in MECH_BASE:
            assert grads2[i][k] == pytest.approx(grads[i][k], rel=1e-10), (i, k)
    ts.mech_end()
    ts.close()
