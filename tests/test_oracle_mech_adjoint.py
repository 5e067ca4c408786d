"""The oracle's frozen mechanics window and its adjoint.

With the reference's parameters a window step is exactly cell_interaction +
pos_copy.  The adjoint (the derivative of every force-law branch, the wall
and the hand-written reverse of tri_bend_1/2) is pinned by central finite
differences of the window; the membrane is given curvature and the FIX
anchors are stretched so that every parameter moves the loss well above
rounding.  The loss is only piecewise smooth (adhesion has kinks at contact
and at 1.2 sigma), so the differences use small steps."""
import numpy as np
import pytest

from epi_amd import configs


def window_tissue(O, seed=8):
    spec = configs.config_tiny(seed, n_free=260, side=20, grid=20, nz=28)
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    t.set_rng(O.mt19937_64_uniform(1, 100000))
    d = O.Driver(0, 0, 0, 0, 0)
    for _ in range(3):
        t.step(d)
    c = t.get_cells()
    st = c["state"]
    rng = np.random.default_rng(seed)
    x, y, z = c["x"].copy(), c["y"].copy(), c["z"].copy()
    z[st == configs.MEMB] += rng.uniform(-0.15, 0.15, int((st == configs.MEMB).sum()))
    z[st == configs.FIX] += 0.6
    t.set_positions(x, y, z)
    return t


def test_window_step_is_cell_interaction(oracle_mod):
    t = window_tissue(oracle_mod)
    x = t.mech_forward(2)
    t.cell_interaction(); t.pos_copy()
    t.cell_interaction(); t.pos_copy()
    c = t.get_cells()
    assert np.array_equal(x[0], c["x"]) and np.array_equal(x[1], c["y"]) and np.array_equal(x[2], c["z"])


@pytest.mark.parametrize("seed", [8, 9])
def test_adjoint_matches_central_differences(oracle_mod, seed):
    O = oracle_mod
    t = window_tissue(O, seed)
    N = 20
    tgt = t.mech_forward(N, eps=0.0105, k_adh=3.15, k_spring=26.25, k_bend=0.525)
    loss, g = t.mech_adjoint(N, tgt)

    def L(**kw):
        d = t.mech_forward(N, **kw) - tgt
        return np.mean(np.sum(d * d, axis=0))

    assert loss == pytest.approx(L(), rel=1e-12)
    base = dict(O.MECH_DEFAULT)
    for name, b in base.items():
        h = b * 1e-5
        fd = (L(**{**base, name: b + h}) - L(**{**base, name: b - h})) / (2 * h)
        assert abs(g[name] - fd) <= 1e-6 * abs(fd), (name, g[name], fd)
