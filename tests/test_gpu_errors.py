"""Failure paths: the HIP library reports the same error as the oracle
(the reference prints and exits) without faulting the device."""
import numpy as np
import pytest

from helpers import sheet_spec

pytestmark = pytest.mark.gpu


def both(O, spec, rng=None):
    from epi_amd import native
    ot = O.OracleTissue(spec["geometry"])
    ot.load(spec)
    gt = native.Tissue(spec["geometry"])
    gt.load(spec)
    if rng is not None:
        ot.set_rng(rng)
        gt.set_rng(rng)
    return ot, gt


def gpu_code(fn):
    from epi_amd import native
    try:
        fn()
    except native.EpiError as e:
        return e.code
    return 0


def test_strong_interaction(oracle_mod):
    free = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=1.0), dict(state=0, x=8.9, y=8.0, z=9.0, agek=1.0)]
    ot, gt = both(oracle_mod, sheet_spec(free))
    ot.cell_interaction()
    assert ot.error() == 2

    def run():
        gt.cell_interaction()
        gt.set.synchronize()
    assert gpu_code(run) == 2


def test_no_dermis(oracle_mod):
    free = [dict(state=4, x=8.0, y=8.0, z=12.0, ageb=1.0)]
    ot, gt = both(oracle_mod, sheet_spec(free))
    ot.cell_state_renew()
    assert ot.error() == 3

    def run():
        gt.cell_state_renew()
        gt.set.synchronize()
    assert gpu_code(run) == 3


def test_remove_alias(oracle_mod):
    free = [dict(state=1, x=3.0, y=3.0, z=9.0, agek=31.5), dict(state=1, x=11.0, y=11.0, z=9.0, agek=31.5)]
    ot, gt = both(oracle_mod, sheet_spec(free))
    ot.cell_state_renew()
    assert ot.error() == 5

    def run():
        gt.cell_state_renew()
        gt.set.synchronize()
    assert gpu_code(run) == 5


def test_rng_exhausted(oracle_mod):
    free = [dict(state=4, x=8.0, y=8.0, z=3.4, ageb=5.0)]
    ot, gt = both(oracle_mod, sheet_spec(free), rng=np.array([0.5]))
    for t in (ot, gt):
        t.cell_state_renew()
    assert ot.error() == 0
    assert gpu_code(lambda: gt.set.synchronize()) == 0
    ot.cell_state_renew()
    assert ot.error() == 6

    def run():
        gt.cell_state_renew()
        gt.set.synchronize()
    assert gpu_code(run) == 6


def test_out_of_grid(oracle_mod):
    free = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=1.0)]
    spec = sheet_spec(free)
    spec["cells"]["z"][-1] = 40.0          # above LZ = 16
    from epi_amd import native
    gt = native.Tissue(spec["geometry"])
    with pytest.raises(native.EpiError) as e:
        gt.load(spec)
    assert e.value.code == 4
    ot = oracle_mod.OracleTissue(spec["geometry"])
    with pytest.raises(RuntimeError):
        ot.load(spec)
    assert ot.error() == 4


def test_bucket_capacity(oracle_mod):
    free = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=1.0)]
    spec = sheet_spec(free, nmx=16, nmy=16)                 # 4 membrane particles per bucket
    from epi_amd import native
    gt = native.Tissue(spec["geometry"], bucket_cap=2)
    with pytest.raises(native.EpiError) as e:
        gt.load(spec)
    assert e.value.code == 1
