"""Launch-shape autotuner for the grid kernels of the epi5_MT hot path.

The shapes it chooses (see epi_set_tuning in include/epi_amd.h):

  k_ext_stim   z planes per tile (1, 2, 4, 8, 16) x 64-lane rows per block (1, 2, 4),
               then its tile list in one run per XCD or not (b_xcd 0, 1)
  k_atp        the same shapes
  k_ca_cells   lanes per cell (4, 8, 16)
  k_ca_cells + k_atp   XCD-contiguous workgroup order (0, 1); in fused mode,
               the fused launch's (its ATP tiles in one run per XCD)
  Ca solve     ATP of a sub-step and cells of the next in one launch (1; 2 =
               only solves of a few tissues, when the whole set runs faster
               unfused), with its own lanes per cell and z planes per thread
  k_map_stamp  lanes per cell (1, 4, 16)
  k_sweep      32 lanes per cell or one wave per cell (sweep_wave 0, 1)
  whole steps  where the side stream forks (side_fork 0, 1)

With CA_FUSE 2 the fused small solves use CA_GROUP_SMALL lanes per cell
(default 8, the best width on a latency-bound single-tissue solve).

A launch shape changes timing only: every grid point is computed on its own
and the Ca cell kernel sums each list in list order whatever its group size,
so the tuner can time candidates on a throwaway copy of the workload (which
it advances) and hand the winners to the real run.  Each candidate is timed
with HIP events on the library's stream (TissueSet.profile).
"""
from __future__ import annotations

import itertools
import json
import statistics
import time

STENCIL_SHAPES = list(itertools.product((1, 2, 4, 8, 16), (1, 2, 4)))
CA_GROUPS = (4, 8, 16)
DEFAULT = {"b_zb": 4, "b_rows": 4, "atp_zb": 4, "atp_rows": 4, "ca_group": 16, "xcd": 0, "ca_fuse": 0, "map_lanes": 1,
           "b_xcd": 0, "sweep_wave": 0, "side_fork": 0}


def _avg_us(ts, kernel):
    prof = ts.profile()
    if kernel not in prof:
        return float("inf")
    ms, cnt = prof[kernel]
    return 1e3 * ms / cnt


def _time(ts, kernel, run):
    """Mean launch time of `kernel` (or the sum over a list of kernels)."""
    kernels = [kernel] if isinstance(kernel, str) else list(kernel)
    ts.synchronize()
    ts.profile_reset()
    ts.profile_sample(1)
    ts.profile_enable(kernels)
    run()
    ts.synchronize()
    us = sum(_avg_us(ts, k) for k in kernels)
    ts.profile_enable([])
    return us


def tune(ts, drivers, warm_steps=3, b_reps=20, ca_sub=40, log=None, step_reps=7):
    """Choose launch shapes on `ts`, a loaded TissueSet that may be advanced.

    Returns (tuning dict, report dict with every candidate's mean launch time
    in microseconds)."""
    t0 = time.perf_counter()
    if warm_steps:
        ts.step(drivers, warm_steps)
    zz = ts.calc_zzmax()
    report = {"k_ext_stim": {}, "k_atp": {}, "k_ca_cells": {}, "xcd": {}, "ca_fuse": {}, "k_map_stamp": {}}
    best = dict(DEFAULT)
    ts.set_tuning(best)

    def run_map():
        for _ in range(b_reps):
            ts.setup_map_lat()

    for ml in (1, 4, 16):
        ts.set_tuning({"map_lanes": ml})
        report["k_map_stamp"][f"lanes{ml}"] = _time(ts, "k_map_stamp", run_map)
    best["map_lanes"] = int(min(report["k_map_stamp"], key=report["k_map_stamp"].get)[5:])
    ts.set_tuning(best)

    def run_lists():
        for _ in range(b_reps):
            ts.connect_cell()

    report["k_sweep"] = {}
    for w in (0, 1):
        ts.set_tuning({"sweep_wave": w})
        report["k_sweep"][f"wave{w}"] = _time(ts, "k_sweep_wave" if w else "k_sweep", run_lists)
    best["sweep_wave"] = int(min(report["k_sweep"], key=report["k_sweep"].get)[4:])
    ts.set_tuning(best)

    def run_b():
        for _ in range(b_reps):
            ts.calc_ext_stim(zz)

    for zb, rows in STENCIL_SHAPES:
        ts.set_tuning({"b_zb": zb, "b_rows": rows})
        ts.calc_ext_stim(zz)
        report["k_ext_stim"][f"zb{zb}_rows{rows}"] = _time(ts, "k_ext_stim", run_b)
    key = min(report["k_ext_stim"], key=report["k_ext_stim"].get)
    best["b_zb"], best["b_rows"] = (int(v) for v in key[2:].split("_rows"))
    ts.set_tuning(best)
    report["b_xcd"] = {}
    for x in (0, 1):
        ts.set_tuning({"b_xcd": x})
        report["b_xcd"][f"xcd{x}"] = _time(ts, "k_ext_stim", run_b)
    best["b_xcd"] = int(min(report["b_xcd"], key=report["b_xcd"].get)[3:])
    ts.set_tuning(best)

    def run_ca():
        ts.calc_ca2p_n(zz, ca_sub)

    for zb, rows in STENCIL_SHAPES:
        ts.set_tuning({"atp_zb": zb, "atp_rows": rows})
        report["k_atp"][f"zb{zb}_rows{rows}"] = _time(ts, "k_atp", run_ca)
    key = min(report["k_atp"], key=report["k_atp"].get)
    best["atp_zb"], best["atp_rows"] = (int(v) for v in key[2:].split("_rows"))
    ts.set_tuning(best)

    for g in CA_GROUPS:
        ts.set_tuning({"ca_group": g})
        report["k_ca_cells"][f"group{g}"] = _time(ts, "k_ca_cells", run_ca)
    key = min(report["k_ca_cells"], key=report["k_ca_cells"].get)
    best["ca_group"] = int(key[5:])
    ts.set_tuning(best)

    report["xcd"] = {}
    for x in (0, 1):
        ts.set_tuning({"xcd": x})
        report["xcd"][f"xcd{x}"] = _time(ts, ("k_ca_cells", "k_atp"), run_ca)
    best["xcd"] = int(min(report["xcd"], key=report["xcd"].get)[3:])
    ts.set_tuning(best)

    # per sub-step: the two kernels against one fused launch, whose lanes per
    # cell and z planes per thread of its grid part are chosen for it
    # (1 or 2 planes: the fused grid part is latency-bound, and 4 or more
    # planes per thread take the column-rolling form meant for large tiles)
    report["ca_fuse"] = {"unfused": report["xcd"][f"xcd{best['xcd']}"]}
    for g in CA_GROUPS:
        for zb in (1, 2):
            ts.set_tuning({"ca_fuse": 1, "ca_group": g, "atp_zb": zb})
            report["ca_fuse"][f"fused_group{g}_zb{zb}"] = _time(ts, "k_ca_fused", run_ca)
    key = min(report["ca_fuse"], key=report["ca_fuse"].get)
    if key != "unfused":
        best["ca_fuse"] = 1
        best["ca_group"], best["atp_zb"] = (int(v) for v in key[len("fused_group"):].split("_zb"))
        # the fused launch's own XCD order (its ATP tiles in one run per XCD)
        ts.set_tuning(best)
        report["xcd_fused"] = {}
        for x in (0, 1):
            ts.set_tuning({"xcd": x})
            report["xcd_fused"][f"xcd{x}"] = _time(ts, "k_ca_fused", run_ca)
        best["xcd"] = int(min(report["xcd_fused"], key=report["xcd_fused"].get)[3:])
    else:
        # the whole set solves faster unfused; a gated solve of a few
        # tissues is latency-bound and still fuses (ca_fuse 2)
        best["ca_fuse"] = 2
    ts.set_tuning(best)

    # where the side stream forks: a property of whole steps (what runs beside
    # what), so whole steps are timed, the median of single steps (a step
    # that an sw-gated solve joins is an outlier); the late fork must win by
    # 2 % to be taken
    report["side_fork"] = {}
    if drivers:
        for f in (0, 1):
            ts.set_tuning({"side_fork": f})
            ts.step(drivers, 1)
            us = []
            for _ in range(step_reps):
                ts.synchronize()
                s0 = time.perf_counter()
                ts.step(drivers, 1)
                ts.synchronize()
                us.append(1e6 * (time.perf_counter() - s0))
            report["side_fork"][f"fork{f}"] = statistics.median(us)
        late = report["side_fork"]["fork1"] < 0.98 * report["side_fork"]["fork0"]
        best["side_fork"] = 1 if late else 0
    ts.set_tuning(best)
    report["seconds"] = time.perf_counter() - t0
    if log is not None:
        print("autotune " + json.dumps({"choice": best, "us": report}), file=log, flush=True)
    return best, report
