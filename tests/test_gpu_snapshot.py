"""Asynchronous snapshots (epi_snapshot_begin / epi_snapshot_wait): the
per-cell record CellManager::output writes (cellmanager.cpp:451-538), taken
behind the queued work while later steps run (proc.cpp:139-141 takes one
every CUT steps).  Checked bit for bit against a synchronous download of the
same moment, with the touch flag recomputed on the host from the lists
(MI355X only)."""
import numpy as np
import pytest

from epi_amd import configs

pytestmark = pytest.mark.gpu

F64 = ("radius", "ageb", "agek", "ca2p", "x", "y", "z", "ca2p_avg", "ex_fat", "in_fat", "spr_nat_len")
I32 = ("state", "rest_div_times", "pair", "fix_origin")


def touch_of(c):
    """output's flag: an ALIVE cell with a DEAD cell in its list"""
    st, ptr, conn = c["state"], c["conn_ptr"], c["conn"]
    out = np.zeros(st.size, dtype=np.uint8)
    for i in np.nonzero(st == configs.ALIVE)[0]:
        if np.any(st[conn[ptr[i]:ptr[i + 1]]] == configs.DEAD):
            out[i] = 1
    return out


def same_record(snap, c):
    assert snap["n"] == c["x"].size
    for k in F64:
        assert np.array_equal(snap[k], c[k]), k
    for k in I32:
        assert np.array_equal(snap[k].astype(np.int64), c[k].astype(np.int64)), k
    assert np.array_equal(snap["touch"], touch_of(c))


def test_snapshot_is_the_state_at_begin_while_steps_run():
    """begin, 15 more steps, then wait: the record is the state at begin;
    the second slot, then the first one again (reused), likewise"""
    from epi_amd import native
    spec = configs.config_tiny(31, n_free=300, side=16, grid=16, n_layers=3, dead_top=True, nz=40)
    t = native.TissueSet([spec["geometry"]])
    t.load(0, spec, True)
    t.set_rng(0, native.mt19937_64_uniform(31, 400000))
    d = [native.Driver(0, 0, 0, 20, 0)]
    print("snapshot: tiny tissue 300 cells, slots 0, 1, 0", flush=True)
    t.step(d, 10)
    for slot in (0, 1, 0):
        ref = t.get_cells(0)
        t.snapshot_begin(0, slot)
        t.step(d, 15)
        snap = t.snapshot_wait(slot)
        assert snap["tissue"] == 0
        same_record(snap, ref)
        assert np.any(touch_of(ref)) or slot == 0
    assert t.status(0)[0].err == 0
    t.close()


def test_snapshot_of_one_tissue_of_a_batch_and_two_in_flight():
    """two slots in flight at once, for two tissues of one context"""
    from epi_amd import native
    specs = [configs.config_tiny(41, n_free=200, side=16, grid=16, n_layers=3, nz=32),
             configs.config_tiny(42, n_free=240, side=16, grid=16, n_layers=3, nz=32)]
    t = native.TissueSet([s["geometry"] for s in specs])
    for i, s in enumerate(specs):
        t.load(i, s, True)
        t.set_rng(i, native.mt19937_64_uniform(41 + i, 300000))
    d = [native.Driver(0, 0, 0, 0, 0) for _ in specs]
    print("snapshot: two tissues, both slots in flight", flush=True)
    t.step(d, 5)
    refs = [t.get_cells(0), t.get_cells(1)]
    t.snapshot_begin(1, 0)
    t.snapshot_begin(0, 1)
    t.step(d, 10)
    b, a = t.snapshot_wait(0), t.snapshot_wait(1)
    assert (a["tissue"], b["tissue"]) == (0, 1)
    same_record(a, refs[0])
    same_record(b, refs[1])
    # a slot can be read again until it is reused
    same_record(t.snapshot_wait(0), refs[1])
    t.close()


def test_wait_on_an_empty_slot_and_a_bad_slot_are_errors():
    from epi_amd import native
    spec = configs.config_tiny(43, n_free=100, side=16, grid=16, nz=32)
    t = native.TissueSet([spec["geometry"]])
    t.load(0, spec, True)
    with pytest.raises(native.EpiError):
        t.snapshot_wait(1)
    with pytest.raises(native.EpiError):
        t.snapshot_begin(0, 2)
    t.close()
