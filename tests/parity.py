"""Comparison helpers shared by the parity tests (oracle vs HIP)."""
from __future__ import annotations

import numpy as np

INT_FIELDS = ("state", "fix_origin", "rest_div_times", "pair", "dermis", "is_malignant",
              "uid", "lat", "conn_ptr", "conn", "memb_adj")
FP_FIELDS = ("x", "y", "z", "xn", "yn", "zn", "radius", "ca2p", "ca2p_next", "ip3", "ip3_next",
             "ca2p_avg", "ex_inert", "agek", "ageb", "ex_fat", "in_fat", "spr_nat_len",
             "diff_u", "gj")


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    den = np.maximum(np.abs(b), 1e-300)
    d = np.abs(a - b)
    # relative where the reference is not tiny, absolute otherwise
    r = np.where(np.abs(b) > 1e-12, d / den, d)
    return float(np.max(r))


def compare_cells(gpu, ref, fields=None, skip=()):
    """Returns {field: mismatch count (ints) or max rel err (floats)}."""
    out = {}
    if gpu["n"] != ref["n"]:
        out["n"] = (gpu["n"], ref["n"])
        return out
    for k in INT_FIELDS:
        if k in skip or (fields and k not in fields):
            continue
        a, b = np.asarray(gpu[k]), np.asarray(ref[k])
        if a.shape != b.shape:
            out[k] = ("shape", a.shape, b.shape)
        else:
            out[k] = int(np.count_nonzero(a != b))
    for k in FP_FIELDS:
        if k in skip or (fields and k not in fields):
            continue
        a, b = gpu[k], ref[k]
        if np.shape(a) != np.shape(b):
            out[k] = ("shape", np.shape(a), np.shape(b))
        else:
            out[k] = rel_err(a, b)
    return out


def compare_fields(gpu, ref, names=("atp", "atp_next", "b", "b_next", "cmap1", "cmap2")):
    out = {}
    for k in names:
        a, b = gpu[k], ref[k]
        if k.startswith("cmap"):
            out[k] = int(np.count_nonzero(a != b))
        else:
            out[k] = rel_err(a, b)
    return out


def worst(d):
    bad = {}
    for k, v in d.items():
        if isinstance(v, tuple):
            bad[k] = v
        elif isinstance(v, int) and v != 0:
            bad[k] = v
        elif isinstance(v, float) and v > 0:
            bad[k] = v
    return bad
