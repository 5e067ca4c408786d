"""The two adjoint windows behind torch.autograd.Function (test harness only;
the library itself takes no torch types), checked with
torch.autograd.gradcheck: the backward pass (the library's hand-written
adjoint, epi_rd_adjoint / epi_mech_adjoint) against torch's central
differences of the forward pass (epi_rd_forward / epi_mech_forward).

The tensors are CPU float64 parameters; the windows themselves run on the
GPU through the C ABI (MI355X only)."""
import numpy as np
import pytest

from epi_amd import configs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
AFM = (configs.ALIVE, configs.FIX, configs.MUSUME)


class RDWindowLoss(torch.autograd.Function):
    """loss = mean over AFM cells of (ca2p after n_sub sub-steps - target)^2
    as a function of (du, dp, gj0, kpa) (ca2p.cpp:392-399 run n_sub times)"""

    @staticmethod
    def forward(ctx, du, dp, gj0, kpa, tissue, n_sub, every, target, afm):
        p = dict(du=float(du), dp=float(dp), gj0=float(gj0), kpa=kpa.detach().numpy().copy())
        tissue.set.rd_forward([p], n_sub)
        u = tissue.set.rd_get_ca2p(0)
        ctx.window = (tissue, n_sub, every, target, p)
        return torch.tensor(np.mean((u[afm] - target[afm]) ** 2), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        tissue, n_sub, every, target, p = ctx.window
        _, grads = tissue.set.rd_adjoint([p], n_sub, every, [target])
        g = grads[0]
        f = lambda v: gout * torch.as_tensor(v, dtype=torch.float64)
        return f(g["du"]), f(g["dp"]), f(g["gj0"]), f(g["kpa"]), None, None, None, None, None


class MechWindowLoss(torch.autograd.Function):
    """loss = mean over cells of |x after n_steps - target|^2 as a function of
    (eps, k_adh, k_spring, k_bend) (cell_interaction + pos_copy run n_steps
    times)"""

    @staticmethod
    def forward(ctx, eps, k_adh, k_spring, k_bend, tissue, n_steps, every, target):
        p = dict(eps=float(eps), k_adh=float(k_adh), k_spring=float(k_spring), k_bend=float(k_bend))
        tissue.set.mech_forward([p], n_steps)
        d = tissue.set.mech_get_positions(0) - target
        ctx.window = (tissue, n_steps, every, target, p)
        return torch.tensor(np.mean(np.sum(d * d, axis=0)), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        tissue, n_steps, every, target, p = ctx.window
        _, grads = tissue.set.mech_adjoint([p], n_steps, every, [target])
        g = grads[0]
        f = lambda k: gout * torch.tensor(g[k], dtype=torch.float64)
        return f("eps"), f("k_adh"), f("k_spring"), f("k_bend"), None, None, None, None


def _tissue(seed, steps=3):
    from epi_amd import native
    spec = configs.config_tiny(seed, n_free=220, side=20, grid=20, nz=28)
    t = native.Tissue(spec["geometry"])
    t.load(spec)
    t.set_rng(native.mt19937_64_uniform(1, 100000))
    t.step(native.Driver(0, 0, 0, 0, 0), steps)
    assert t.error() == 0
    return t


def test_rd_window_gradcheck():
    t = _tissue(8)
    zz = t.calc_zzmax()
    st = t.get_cells()["state"]
    afm = np.isin(st, AFM)
    n_sub, every = 12, 5
    rng = np.random.default_rng(5)
    kpa0 = np.where(st == configs.ALIVE, 4.5, 6.0) * (1 + 0.05 * rng.uniform(-1, 1, st.size))
    t.set.rd_begin(zz)
    t.set.rd_forward([dict(du=0.0105, dp=0.095, gj0=0.95, kpa=kpa0 * 1.05)], n_sub)
    target = t.set.rd_get_ca2p(0)
    # kpa of a few AFM cells is differentiated (every entry costs two forward
    # windows in gradcheck); the others enter as constants through a mask
    idx = np.nonzero(afm)[0][:6]
    base = torch.tensor(kpa0, dtype=torch.float64)
    sel = torch.zeros(st.size, dtype=torch.float64)
    sel[idx] = 1.0

    def f(du, dp, gj0, kv):
        kpa = base * (1 - sel) + torch.zeros_like(base).index_put((torch.as_tensor(idx),), kv)
        return RDWindowLoss.apply(du, dp, gj0, kpa, t, n_sub, every, target, afm)

    args = (torch.tensor(0.01, dtype=torch.float64, requires_grad=True),
            torch.tensor(0.1, dtype=torch.float64, requires_grad=True),
            torch.tensor(0.99, dtype=torch.float64, requires_grad=True),
            torch.tensor(kpa0[idx], dtype=torch.float64, requires_grad=True))
    print("gradcheck of the reaction-diffusion window (12 sub-steps)", flush=True)
    # the adjoint sums with FP64 atomics, so two backward passes agree to
    # rounding only (nondet_tol)
    assert torch.autograd.gradcheck(f, args, eps=1e-6, atol=1e-9, rtol=1e-5, nondet_tol=1e-12)
    t.set.rd_end()


def test_mech_window_gradcheck():
    t = _tissue(9)
    c = t.get_cells()
    st = c["state"]
    rng = np.random.default_rng(9)
    x, y, z = c["x"].copy(), c["y"].copy(), c["z"].copy()
    z[st == configs.MEMB] += rng.uniform(-0.15, 0.15, int((st == configs.MEMB).sum()))
    z[st == configs.FIX] += 0.6
    t.set_positions(x, y, z)
    n_steps, every = 10, 4
    t.set.mech_begin()
    t.set.mech_forward([dict(eps=0.0105, k_adh=3.15, k_spring=26.25, k_bend=0.525)], n_steps)
    target = t.set.mech_get_positions(0)
    args = tuple(torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (0.01, 3.0, 25.0, 0.5))
    print("gradcheck of the mechanics window (10 steps)", flush=True)
    # the loss is piecewise smooth (adhesion kinks at contact and 1.2 sigma),
    # hence the small difference step
    assert torch.autograd.gradcheck(lambda *a: MechWindowLoss.apply(*a, t, n_steps, every, target), args,
                                    eps=1e-7, atol=1e-9, rtol=1e-5, nondet_tol=1e-12)
    t.set.mech_end()
