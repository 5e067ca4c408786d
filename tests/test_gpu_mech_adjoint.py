"""The frozen mechanics window and its adjoint on the GPU (MI355X).

Forward: a window step is the plain cell_interaction + pos_copy (bit-exact)
and equals the oracle's window (bit-exact positions).  Reverse: the gradient
agrees with the oracle's reverse pass (1e-8: FP64 atomics reorder sums) and
with central differences of the GPU window (1e-6); it does not depend on the
checkpoint interval; tissues batched in one context get their own
gradients; epi_mech_end restores the positions."""
import numpy as np
import pytest

from epi_amd import configs

pytestmark = pytest.mark.gpu

TARGET_PARAMS = dict(eps=0.0105, k_adh=3.15, k_spring=26.25, k_bend=0.525)


def _perturb(c, seed):
    st = c["state"]
    rng = np.random.default_rng(seed)
    x, y, z = c["x"].copy(), c["y"].copy(), c["z"].copy()
    z[st == configs.MEMB] += rng.uniform(-0.15, 0.15, int((st == configs.MEMB).sum()))
    z[st == configs.FIX] += 0.6
    return x, y, z


def _spec(seed):
    return configs.config_tiny(seed, n_free=260, side=20, grid=20, nz=28)


def _pair(O, seed=8):
    from epi_amd import native
    spec = _spec(seed)
    draws = native.mt19937_64_uniform(1, 100000)
    o = O.OracleTissue(spec["geometry"])
    o.load(spec)
    o.set_rng(draws)
    t = native.Tissue(spec["geometry"])
    t.load(spec)
    t.set_rng(draws)
    d = O.Driver(0, 0, 0, 0, 0)
    for _ in range(3):
        o.step(d)
    t.step(native.Driver(0, 0, 0, 0, 0), 3)
    x, y, z = _perturb(o.get_cells(), seed)
    o.set_positions(x, y, z)
    t.set_positions(x, y, z)
    return o, t


def test_window_forward(oracle_mod):
    o, t = _pair(oracle_mod)
    print("mech window forward: tiny tissue, 12 steps", flush=True)
    t.set.mech_begin()
    t.set.mech_forward([{}], 12)
    xg = t.set.mech_get_positions(0)
    t.set.mech_forward([TARGET_PARAMS], 12)
    xg2 = t.set.mech_get_positions(0)
    t.set.mech_end()
    assert np.array_equal(xg, o.mech_forward(12))
    assert np.array_equal(xg2, o.mech_forward(12, **TARGET_PARAMS))
    # the window step is the plain step without the periodic fix
    for _ in range(12):
        t.cell_interaction()
        t.pos_copy()
    c = t.get_cells()
    assert np.array_equal(xg[0], c["x"]) and np.array_equal(xg[2], c["z"])


@pytest.mark.parametrize("n_steps,every", [(20, 20), (30, 7)])
def test_adjoint_against_oracle_and_differences(oracle_mod, n_steps, every):
    o, t = _pair(oracle_mod)
    tgt = o.mech_forward(n_steps, **TARGET_PARAMS)
    lo, go = o.mech_adjoint(n_steps, tgt)
    print(f"mech adjoint: {n_steps} steps, checkpoint every {every}", flush=True)
    t.set.mech_begin()
    loss, grads = t.set.mech_adjoint([{}], n_steps, every, [tgt])
    g = grads[0]
    assert loss[0] == pytest.approx(lo, rel=1e-12)
    for k in go:
        assert g[k] == pytest.approx(go[k], rel=1e-8), k
    _, g1 = t.set.mech_adjoint([{}], n_steps, 1, [tgt])
    for k in go:
        assert g1[0][k] == pytest.approx(g[k], rel=1e-10), k
    base = dict(eps=0.01, k_adh=3.0, k_spring=25.0, k_bend=0.5)

    def L(**kw):
        t.set.mech_forward([{**base, **kw}], n_steps)
        d = t.set.mech_get_positions(0) - tgt
        return np.mean(np.sum(d * d, axis=0))

    for name, b in base.items():
        h = b * 1e-5
        fd = (L(**{name: b + h}) - L(**{name: b - h})) / (2 * h)
        assert abs(g[name] - fd) <= 1e-6 * abs(fd), (name, g[name], fd)
    before = t.get_cells()
    t.set.mech_end()
    after = t.get_cells()
    # end restores the window's start, not the last forward
    assert not np.array_equal(before["x"], after["x"]) or n_steps == 0


def test_batched_windows_are_independent(oracle_mod):
    from epi_amd import native
    seeds = [8, 9, 10]
    specs = [_spec(s) for s in seeds]
    ts = native.TissueSet([s["geometry"] for s in specs])
    singles = []
    draws = native.mt19937_64_uniform(1, 100000)
    for i, s in enumerate(specs):
        ts.load(i, s)
        ts.set_rng(i, draws)
    ts.step([native.Driver(0, 0, 0, 0, 0) for _ in specs], 3)
    tgts, params = [], []
    for i, s in enumerate(seeds):
        x, y, z = _perturb(ts.get_cells(i), s)
        ts.set_positions(i, x, y, z)
        o = oracle_mod.OracleTissue(specs[i]["geometry"])
        o.load(specs[i])
        o.set_rng(draws)
        d = oracle_mod.Driver(0, 0, 0, 0, 0)
        for _ in range(3):
            o.step(d)
        o.set_positions(x, y, z)
        tgts.append(o.mech_forward(15, **TARGET_PARAMS))
        params.append(dict(eps=0.01 * (1 + 0.01 * i), k_bend=0.5 * (1 - 0.01 * i)))
        singles.append(o.mech_adjoint(15, tgts[-1], **params[-1]))
    ts.mech_begin()
    loss, grads = ts.mech_adjoint(params, 15, 5, tgts)
    ts.mech_end()
    for i in range(len(seeds)):
        lo, go = singles[i]
        assert loss[i] == pytest.approx(lo, rel=1e-12)
        for k in go:
            assert grads[i][k] == pytest.approx(go[k], rel=1e-8), (i, k)
