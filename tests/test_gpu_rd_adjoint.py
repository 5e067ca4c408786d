"""The frozen reaction-diffusion window and its adjoint on the GPU (MI355X).

Forward: with the reference's parameters the window is calc_ca2p's loop
(bit-exact); with other parameters it agrees with the oracle's window within
1e-10.  Reverse: the gradient agrees with the oracle's reverse pass (FP64
atomics reorder sums: 1e-8) and with central finite differences of the GPU
window itself (1e-6), for every checkpoint interval, and epi_rd_end leaves
the tissue as it was."""
import numpy as np
import pytest

from epi_amd import configs

pytestmark = pytest.mark.gpu

AFM = (configs.ALIVE, configs.FIX, configs.MUSUME)


def _pair(O, spec, steps=3):
    from epi_amd import native
    g = spec["geometry"]
    draws = native.mt19937_64_uniform(1, 100000)
    o = O.OracleTissue(g)
    o.load(spec)
    o.set_rng(draws)
    t = native.Tissue(g)
    t.load(spec)
    t.set_rng(draws)
    if steps:
        d = O.Driver(0, 0, 0, 0, 0)
        for _ in range(steps):
            o.step(d)
        t.step(native.Driver(0, 0, 0, 0, 0), steps)
    assert o.error() == 0 and t.error() == 0
    return o, t


def _spec(seed=8):
    return configs.config_tiny(seed, n_free=220, side=20, grid=20, nz=28)


def test_window_forward(oracle_mod):
    spec = configs.config_tiny(7, n_free=200, nz=24)
    o, t = _pair(oracle_mod, spec, steps=0)
    zz = t.calc_zzmax()
    st = t.get_cells()["state"]
    afm = np.isin(st, AFM)
    print("rd window forward: tiny tissue, 12 sub-steps", flush=True)
    t.set.rd_begin(zz)
    t.set.rd_forward([{}], 12)
    u = t.set.rd_get_ca2p(0)
    t.set.rd_end()
    t.set.calc_ca2p_n(zz, 12)
    ref = t.get_cells()["ca2p"]
    assert np.array_equal(u[afm], ref[afm])
    # perturbed parameters against the oracle's window
    o2, t2 = _pair(oracle_mod, _spec())
    zz = t2.calc_zzmax()
    st = t2.get_cells()["state"]
    kpa = np.where(st == configs.ALIVE, 4.7, 6.1)
    prm = dict(du=0.0107, dp=0.093, gj0=0.95, kpa=kpa)
    t2.set.rd_begin(zz)
    t2.set.rd_forward([prm], 30)
    ug = t2.set.rd_get_ca2p(0)
    t2.set.rd_end()
    uo = o2.rd_forward(zz, 30, **prm)
    afm = np.isin(st, AFM)
    rel = np.max(np.abs(ug[afm] - uo[afm]) / np.abs(uo[afm]))
    assert rel <= 1e-10, rel


def test_window_end_restores_the_tissue(oracle_mod):
    o, t = _pair(oracle_mod, _spec())
    before = t.get_cells()
    fb = t.get_fields()
    zz = t.calc_zzmax()
    t.set.rd_begin(zz)
    t.set.rd_forward([dict(du=0.02, gj0=0.5)], 15)
    tgt = np.zeros(before["x"].size)
    t.set.rd_adjoint([{}], 15, 4, [tgt])
    t.set.rd_end()
    after = t.get_cells()
    fa = t.get_fields()
    for k in ("ca2p", "ca2p_next", "ip3", "ip3_next", "ex_inert", "ca2p_avg", "gj", "x", "conn"):
        assert np.array_equal(before[k], after[k]), k
    for k in ("atp", "atp_next", "b", "b_next"):
        assert np.array_equal(fb[k], fa[k]), k
    # the tissue then steps exactly like the oracle again
    d = oracle_mod.Driver(0, 0, 0, 0, 0)
    for _ in range(3):
        o.step(d)
    from epi_amd import native
    t.step(native.Driver(0, 0, 0, 0, 0), 3)
    assert np.array_equal(t.get_cells()["x"], o.get_cells()["x"])


@pytest.mark.parametrize("n_sub,every", [(20, 20), (40, 7)])
def test_adjoint_against_oracle_and_differences(oracle_mod, n_sub, every):
    o, t = _pair(oracle_mod, _spec())
    zz = t.calc_zzmax()
    st = t.get_cells()["state"]
    n = st.size
    afm = np.isin(st, AFM)
    rng = np.random.default_rng(3)
    kpa = np.where(st == configs.ALIVE, 4.5, 6.0) * (1 + 0.05 * rng.uniform(-1, 1, n))
    tgt = o.rd_forward(zz, n_sub, du=0.0105, dp=0.095, gj0=0.95, kpa=kpa * 1.05)
    base = dict(du=0.01, dp=0.1, gj0=0.99, kpa=kpa)
    lo, go = o.rd_adjoint(zz, n_sub, tgt, **base)
    print(f"rd adjoint: {n_sub} sub-steps, checkpoint every {every}", flush=True)
    t.set.rd_begin(zz)
    loss, grads = t.set.rd_adjoint([base], n_sub, every, [tgt])
    g = grads[0]
    assert loss[0] == pytest.approx(lo, rel=1e-10)
    for k in ("du", "dp", "gj0"):
        assert g[k] == pytest.approx(go[k], rel=1e-8), k
    sel = np.abs(go["kpa"]) > 1e-3 * np.max(np.abs(go["kpa"]))
    assert np.max(np.abs(g["kpa"][sel] - go["kpa"][sel]) / np.abs(go["kpa"][sel])) <= 1e-8
    assert np.all(g["kpa"][~afm] == 0)
    # neither the checkpoint interval nor the lanes per cell of the reverse
    # kernel changes the result beyond the atomics' summation order
    _, g1 = t.set.rd_adjoint([base], n_sub, 1, [tgt])
    for k in ("du", "dp", "gj0"):
        assert g1[0][k] == pytest.approx(g[k], rel=1e-12), k
    t.set.set_tuning({"rd_group": 16})
    _, g16 = t.set.rd_adjoint([base], n_sub, every, [tgt])
    t.set.set_tuning({"rd_group": 8})
    for k in ("du", "dp", "gj0"):
        assert g16[0][k] == pytest.approx(g[k], rel=1e-12), k
    assert np.allclose(g16[0]["kpa"], g["kpa"], rtol=1e-12, atol=0)

    def L(**kw):
        p = dict(base)
        p.update(kw)
        t.set.rd_forward([p], n_sub)
        u = t.set.rd_get_ca2p(0)
        return np.mean((u[afm] - tgt[afm]) ** 2)

    for name, b in (("du", 0.01), ("dp", 0.1), ("gj0", 0.99)):
        h = b * 1e-5
        fd = (L(**{name: b + h}) - L(**{name: b - h})) / (2 * h)
        assert abs(g[name] - fd) <= 1e-6 * abs(fd), (name, g[name], fd)
    for i in rng.choice(np.nonzero(afm)[0], 3, replace=False):
        h = kpa[i] * 1e-5
        kp, km = kpa.copy(), kpa.copy()
        kp[i] += h
        km[i] -= h
        fd = (L(kpa=kp) - L(kpa=km)) / (2 * h)
        assert abs(g["kpa"][i] - fd) <= 1e-6 * abs(fd), (i, g["kpa"][i], fd)
    t.set.rd_end()


def test_window_after_a_status_read_back(oracle_mod):
    """A window's list of ATP tiles takes iz_bound from the zzmax the host
    holds: a forward window (which ends with a status read-back) followed by
    an adjoint and another forward must still give the oracle's loss and
    values."""
    o, t = _pair(oracle_mod, _spec())
    zz = t.calc_zzmax()
    st = t.get_cells()["state"]
    afm = np.isin(st, AFM)
    pert = dict(du=0.0105, dp=0.095, gj0=0.95)
    n_sub = 30
    tgt = o.rd_forward(zz, n_sub, **pert)
    lo, _ = o.rd_adjoint(zz, n_sub, tgt)
    t.set.rd_begin(zz)
    t.set.rd_forward([pert], n_sub)
    t.set.synchronize()
    loss, _ = t.set.rd_adjoint([{}], n_sub, 10, [tgt])
    assert loss[0] == pytest.approx(lo, rel=1e-10)
    t.set.rd_forward([pert], n_sub)
    u = t.set.rd_get_ca2p(0)
    assert np.max(np.abs(u[afm] - tgt[afm]) / np.abs(tgt[afm])) <= 1e-10
    t.set.rd_end()
