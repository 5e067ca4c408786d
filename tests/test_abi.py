"""The C-ABI library loads and exports every entry point include/epi_amd.h
declares; host-only entry points run without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "epi_amd.h")).read()
    return sorted(set(re.findall(r"\b(epi_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_call_surface():
    syms = declared_symbols()
    for s in ("epi_connect_cell", "epi_cell_interaction", "epi_cell_state_renew",
              "epi_calc_ca2p", "epi_initialize_sc", "epi_setup_map_lat",
              "epi_calc_ext_stim", "epi_calc_zzmax", "epi_pos_copy",
              "epi_cell_pos_periodic_fix", "epi_step"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from epi_amd import native
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(native.SYMBOLS) <= set(declared_symbols())


def test_host_rng_matches_oracle_stream(oracle_mod):
    from epi_amd import native
    a = native.mt19937_64_uniform(12345, 5000)
    b = oracle_mod.mt19937_64_uniform(12345, 5000)
    assert np.array_equal(a, b)
    assert np.all((a >= 0) & (a < 1))


def test_invalid_geometry_is_rejected_without_gpu():
    from epi_amd import native
    g = native.Geometry(2, 3, 8, 8, 8, 10.0, 10.0, 10.0, 100, 0, 0)   # nmy odd, nmx < 3
    h = ctypes.c_void_p()
    rc = native.load_library().epi_create(1, ctypes.byref(g), ctypes.byref(h))
    assert rc == 7 and not h.value
