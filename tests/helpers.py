"""Small hand-built tissues for kernel-level tests."""
from __future__ import annotations

import numpy as np

from epi_amd import configs


def sheet_spec(free, nmx=6, nmy=6, L=16.0, grid=(16, 16, 16), lz=None, sheet_z=1.0,
               max_cells=None, values=True):
    """A flat nmx x nmy sheet at z = sheet_z plus `free` cells, each a dict
    with state, x, y, z and optional radius/agek/ageb/rest_div_times/pair/...
    Lateral box L x L, vertical box lz (default: L)."""
    lz = lz if lz is not None else L
    sheet = configs.make_sheet(nmx, nmy, L, L, z=sheet_z)
    part = configs._empty_cells(len(free))
    for i, f in enumerate(free):
        part["state"][i] = f["state"]
        part["x"][i], part["y"][i], part["z"][i] = f["x"], f["y"], f["z"]
        part["radius"][i] = f.get("radius", 1.4)
        for k in ("agek", "ageb", "ex_fat", "in_fat", "spr_nat_len", "ca2p", "ca2p_avg", "ip3"):
            if k in f:
                part[k][i] = f[k]
        part["rest_div_times"][i] = f.get("rest_div_times", 0)
        part["fix_origin"][i] = f.get("fix_origin", 0)
        part["pair"][i] = f.get("pair", -1)
    nmemb = nmx * nmy
    for i, f in enumerate(free):
        if f.get("pair", -1) >= 0:
            part["pair"][i] = f["pair"] + nmemb
    cells = configs._concat([sheet, part])
    n = cells["x"].size
    cells["ex_inert"][:] = 0.97
    geo = configs.Geometry(nmx=nmx, nmy=nmy, nx=grid[0], ny=grid[1], nz=grid[2], lx=L, ly=L,
                           lz=lz, max_cells=max_cells or n + 256)
    spec = {"geometry": geo, "nmemb": nmemb, "nder": 0, "cells": cells, "seed": 0,
            "values": None, "fields": None}
    if values:
        spec["values"] = {"ca2p": cells["ca2p"], "ip3": cells["ip3"],
                          "ex_inert": cells["ex_inert"], "ca2p_avg": cells["ca2p_avg"]}
    return spec


def lists(cells):
    p, c = cells["conn_ptr"], cells["conn"]
    return [c[p[i]:p[i + 1]].tolist() for i in range(cells["n"])]


def pdiff(a, b, L):
    d = a - b
    return np.where(d > 0.5 * L, d - L, np.where(d <= -0.5 * L, d + L, d))
