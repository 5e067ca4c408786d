"""The gap-junction carry-over search of the row assembly (gj_carry.h, used
by k_finish), built for the host from the same header: for every entry of a
rebuilt row it must find the previous row's slot of the same neighbour uid
(gj_refresh, cell_connection.cpp:133-152: a kept edge keeps its value, a new
one starts at gj_init) and read no slot outside the previous row -- also for
entries at or past twice the previous row's length, where a search that
wrapped by subtracting the length once read stale slots past the row."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def carry(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("carry") / "libcarry.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", os.path.join(HERE, "native", "gj_carry_check.cpp"), "-o", out],
                   check=True)
    L = C.CDLL(out)
    L.carry_slot.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int64,
                             C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.carry_slot.restype = C.c_int

    def run(e, prev, u):
        prev = np.ascontiguousarray(prev, dtype=np.int64)
        lo, hi = C.c_int(), C.c_int()
        # a row with no previous entries hands the search no array at all
        p = prev.ctypes.data_as(C.POINTER(C.c_int64)) if prev.size else None
        q = L.carry_slot(e, prev.size, p, int(u), C.byref(lo), C.byref(hi))
        return q, lo.value, hi.value
    return run


def test_kept_and_new_edges_on_long_rows(carry):
    rng = np.random.default_rng(7)
    for trial in range(300):
        olen = int(rng.integers(0, 12))
        prev = rng.choice(10_000, size=olen, replace=False)
        # the new row: some kept neighbours (shuffled) plus new ones, up to
        # five times the previous length
        keep = rng.permutation(prev)[:int(rng.integers(0, olen + 1))]
        new = rng.choice(np.arange(10_000, 20_000), size=int(rng.integers(0, 4 * olen + 5)), replace=False)
        row = rng.permutation(np.concatenate([keep, new]))
        for e, u in enumerate(row):
            q, lo, hi = carry(e, prev, u)
            want = int(np.nonzero(prev == u)[0][0]) if u in prev else -1
            assert q == want, (trial, e, olen)
            if olen:
                assert 0 <= lo and hi < olen, (trial, e, olen, lo, hi)
            else:
                assert hi == -1                       # nothing read


def test_search_starts_at_the_entry_modulo_the_previous_length(carry):
    # a kept edge at its old position is found with one read
    prev = np.array([5, 9, 2, 7])
    for e in range(len(prev)):
        q, lo, hi = carry(e, prev, prev[e])
        assert q == e and lo == hi == e
    # an entry far past the end starts at e % olen and wraps inside the row
    q, lo, hi = carry(9, prev, 5)          # 9 % 4 = 1: reads 1, 2, 3, then 0
    assert q == 0 and (lo, hi) == (0, 3)
