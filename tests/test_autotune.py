"""The autotuner's choice logic on a stand-in tissue set whose kernel timings
are a function of the launch shape (no GPU: the tuner only sees
TissueSet.profile)."""
from epi_amd import autotune


class FakeSet:
    def __init__(self, cost):
        self.cost = cost
        self.tu = {}
        self.enabled = []

    # the TissueSet calls the tuner makes
    def step(self, drivers, n):
        pass

    def calc_zzmax(self):
        return [10.0]

    def setup_map_lat(self):
        pass

    def connect_cell(self):
        pass

    def calc_ext_stim(self, zz):
        pass

    def calc_ca2p_n(self, zz, n):
        pass

    def synchronize(self):
        pass

    def set_tuning(self, d):
        self.tu.update(d)

    def profile_reset(self):
        pass

    def profile_sample(self, period, run=5):
        pass

    def profile_enable(self, kernels):
        self.enabled = list(kernels)

    def profile(self):
        return {k: (self.cost(k, dict(self.tu)) / 1e3, 1) for k in self.enabled}


def cost_fused_wins(k, tu):
    if k == "k_ca_fused":
        return 10.0 + abs(tu["ca_group"] - 8) + abs(tu["atp_zb"] - 2)
    if k == "k_ca_cells":
        return 9.0 + (tu["ca_group"] == 16) * -1.0 + tu["xcd"]
    if k == "k_atp":
        return 6.0 + tu["atp_zb"] * 0.1 + tu["atp_rows"] * 0.01
    if k == "k_ext_stim":
        return 5.0 + tu["b_zb"] * 0.1 + abs(tu["b_rows"] - 2) * 0.01 - tu["b_xcd"] * 0.5
    if k == "k_map_stamp":
        return 30.0 if tu["map_lanes"] == 1 else 20.0
    return 1.0


def test_fused_shape_is_tuned_in_fused_mode():
    ts = FakeSet(cost_fused_wins)
    best, rep = autotune.tune(ts, [], warm_steps=0)
    # the unfused cell kernel prefers 16 lanes, the fused launch 8 lanes and
    # 2 planes per thread: the fused launch's own optimum is taken
    assert rep["k_ca_cells"]["group16"] < rep["k_ca_cells"]["group8"]
    assert best["ca_fuse"] == 1 and best["ca_group"] == 8 and best["atp_zb"] == 2
    assert best["b_zb"] == 1 and best["b_rows"] == 2 and best["map_lanes"] == 4 and best["b_xcd"] == 1
    assert ts.tu == best
    assert len([k for k in rep["ca_fuse"] if k.startswith("fused_")]) == 6   # 3 lane counts x 1, 2 planes


def test_unfused_wins_gives_fuse_2():
    def cost(k, tu):
        if k == "k_ca_fused":
            return 100.0
        return cost_fused_wins(k, tu)

    ts = FakeSet(cost)
    best, rep = autotune.tune(ts, [], warm_steps=0)
    assert best["ca_fuse"] == 2
    assert best["ca_group"] == 16 and best["xcd"] == 0
    assert best["atp_zb"] == 1 and best["atp_rows"] == 1
