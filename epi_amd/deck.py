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
    names = ["index", "state", "radius", "ageb", "agek", "ca2p", "x", "y", "z", "ca2p_avg",
             "rest_div_times", "ex_fat", "in_fat", "touch", "spr_nat_len", "pair", "fix_origin"]
    return {k: rows[:, i] for i, k in enumerate(names)}


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
    c["ip3"] = np.zeros(n)           # IP3_init (define.h:171)
    c["ex_inert"] = np.full(n, 0.97)  # ex_inert_init (define.h:168)
    pair = np.asarray(cols["pair"][:n]).astype(np.int32)
    c["pair"] = np.where((pair >= 0) & (pair < n), pair, -1).astype(np.int32)
    nmemb = int(np.count_nonzero(c["state"] == configs.MEMB))
    nder = int(np.count_nonzero(c["state"] == configs.DER))
    return {"geometry": g, "nmemb": nmemb, "nder": nder, "cells": c, "values": None,
            "fields": None, "seed": 0}


def load_npz(path, geometry=None):
    with np.load(path) as z:
        cols = {k: z[k] for k in z.files}
    return spec_from_columns(cols, geometry)
