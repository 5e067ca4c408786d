"""The reference's 17-column cell deck (CellManager::_load_from_file,
cellmanager.cpp:198-323) as a tissue spec.

Columns: index, state, radius, ageb, agek, ca2p, x, y, z, ca2p_avg,
rest_div_times, ex_fat, in_fat, touch, spr_nat_len, pair, fix_origin.
Reading stops at the first BLANK row; rows must come MEMB, then DER, then
the rest.  As in the reference's default start (cman_init without
use_last), cell values are then set by _cman_value_init and both fields
start at zero.
"""
from __future__ import annotations

import numpy as np

from . import configs

BLANK = configs.BLANK


def parse_deck(path):
    rows = np.loadtxt(path, dtype=np.float64, ndmin=2)
    return {k: rows[:, i] for i, k in enumerate(DECK_COLUMNS)}


def spec_from_columns(cols, geometry=None):
    state = np.asarray(cols["state"]).astype(np.int64)
    stop = np.nonzero(state == BLANK)[0]
    n = int(stop[0]) if stop.size else state.size
    g = geometry or configs.default_geometry()
    c = configs._empty_cells(n)
    c["state"] = state[:n].astype(np.uint8)
    for k in ("radius", "ageb", "agek", "ca2p", "x", "y", "z", "ca2p_avg", "ex_fat",
              "in_fat", "spr_nat_len"):
        c[k] = np.asarray(cols[k][:n], dtype=np.float64)
    c["rest_div_times"] = np.asarray(cols["rest_div_times"][:n]).astype(np.int32)
    c["fix_origin"] = np.asarray(cols["fix_origin"][:n]).astype(np.int32)
    # is_malignant = stem_orig_id < MALIG_NUM with MALIG_NUM = 0 (cellmanager.cpp:292,
    # define.h:92); the ageing rates it scales are multiplied by 1.0 (cell_state_renew.cpp:32-60)
    c["is_malignant"] = (c["fix_origin"] < 0).astype(np.uint8)
    c["ip3"] = np.zeros(n)           # IP3_init (define.h:171)
    c["ex_inert"] = np.full(n, 0.97)  # ex_inert_init (define.h:168)
    pair = np.asarray(cols["pair"][:n]).astype(np.int32)
    c["pair"] = np.where((pair >= 0) & (pair < n), pair, -1).astype(np.int32)
    nmemb = int(np.count_nonzero(c["state"] == configs.MEMB))
    nder = int(np.count_nonzero(c["state"] == configs.DER))
    return {"geometry": g, "nmemb": nmemb, "nder": nder, "cells": c, "values": None,
            "fields": None, "seed": 0}


DECK_COLUMNS = ["index", "state", "radius", "ageb", "agek", "ca2p", "x", "y", "z", "ca2p_avg",
                "rest_div_times", "ex_fat", "in_fat", "touch", "spr_nat_len", "pair", "fix_origin"]


def write_deck(spec, path):
    """Write the cells of a spec as a 17-column deck that both loaders read
    back to the same doubles (17 significant digits), ending with a BLANK row."""
    c = spec["cells"]
    n = c["state"].size

    def r(k, i):
        return repr(float(c[k][i]))

    with open(path, "w") as f:
        for i in range(n):
            f.write(f"{i} {int(c['state'][i])} {r('radius', i)} {r('ageb', i)} {r('agek', i)} "
                    f"{r('ca2p', i)} {r('x', i)} {r('y', i)} {r('z', i)} {r('ca2p_avg', i)} "
                    f"{int(c['rest_div_times'][i])} {r('ex_fat', i)} {r('in_fat', i)} 0 "
                    f"{r('spr_nat_len', i)} {int(c['pair'][i])} {int(c['fix_origin'][i])}\n")
        f.write(f"{n} {BLANK} 0 0 0 0 0 0 0 0 0 0 0 0 0 -1 -1\n")


def load_npz(path, geometry=None):
    with np.load(path) as z:
        cols = {k: z[k] for k in z.files}
    return spec_from_columns(cols, geometry)
