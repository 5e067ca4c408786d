"""The oracle's frozen reaction-diffusion window and its adjoint.

The window is calc_ca2p's sub-step with frozen positions, lists, states and
maps; with the reference's parameters it must be exactly calc_ca2p's loop.
The adjoint is pinned by central finite differences of the window itself
(a dropped term, a wrong sign or a transposed index in the reverse pass
moves the gradient far outside the 1e-6 bound)."""
import numpy as np
import pytest

from epi_amd import configs


def _tissue(O, spec, steps=3):
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    t.set_rng(O.mt19937_64_uniform(1, 100000))
    d = O.Driver(0, 0, 0, 0, 0)
    for _ in range(steps):
        t.step(d)
    return t


def test_window_with_reference_parameters_is_calc_ca2p(oracle_mod):
    spec = configs.config_tiny(7, n_free=200, nz=24)
    t = _tissue(oracle_mod, spec, steps=0)   # fresh lists: every gj is gj_init
    zz = t.calc_zzmax()
    u = t.rd_forward(zz, 12)
    st = t.get_cells()["state"]
    t.calc_ca2p_n(zz, 12)
    ref = t.get_cells()["ca2p"]
    afm = (st == configs.ALIVE) | (st == configs.FIX) | (st == configs.MUSUME)
    assert afm.sum() > 50
    assert np.array_equal(u[afm], ref[afm])


@pytest.mark.parametrize("seed,n_sub", [(7, 10), (8, 40)])
def test_adjoint_matches_central_differences(oracle_mod, seed, n_sub):
    spec = configs.config_tiny(seed, n_free=220, side=20, grid=20, nz=28)
    t = _tissue(oracle_mod, spec)
    zz = t.calc_zzmax()
    st = t.get_cells()["state"]
    n = st.size
    afm = (st == configs.ALIVE) | (st == configs.FIX) | (st == configs.MUSUME)
    rng = np.random.default_rng(seed)
    kpa = np.where(st == configs.ALIVE, 4.5, 6.0) * (1 + 0.05 * rng.uniform(-1, 1, n))
    # target from the forward model with parameters perturbed by 5%
    tgt = t.rd_forward(zz, n_sub, du=0.0105, dp=0.095, gj0=0.95, kpa=kpa * 1.05)
    loss, g = t.rd_adjoint(zz, n_sub, tgt, kpa=kpa)

    def L(**kw):
        base = dict(du=0.01, dp=0.1, gj0=0.99, kpa=kpa)
        base.update(kw)
        u = t.rd_forward(zz, n_sub, **base)
        return np.mean((u[afm] - tgt[afm]) ** 2)

    assert loss == pytest.approx(L(), rel=1e-14)
    for name, b in (("du", 0.01), ("dp", 0.1), ("gj0", 0.99)):
        h = b * 1e-5
        fd = (L(**{name: b + h}) - L(**{name: b - h})) / (2 * h)
        assert abs(g[name] - fd) <= 1e-6 * abs(fd), (name, g[name], fd)
    for i in rng.choice(np.nonzero(afm)[0], 4, replace=False):
        h = kpa[i] * 1e-5
        kp, km = kpa.copy(), kpa.copy()
        kp[i] += h
        km[i] -= h
        fd = (L(kpa=kp) - L(kpa=km)) / (2 * h)
        assert abs(g["kpa"][i] - fd) <= 1e-6 * abs(fd), (i, g["kpa"][i], fd)
    # only AFM cells produce IP3 in supra_calc
    assert np.all(g["kpa"][~afm] == 0)
