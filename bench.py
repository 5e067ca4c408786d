#!/usr/bin/env python3
"""Throughput of the epi5_MT per-step hot path on one MI355X.

Workload (default): the plain FP64 large single tissue -- 128x128 membrane
sheet, 64 FIX, 20,000 free cells in 8 layers (top layer DEAD), FP64 grids
128x128x256, division and removal active, seed 2 -- stepped through
proc.cpp:153-172 with the reference's sw-gated Ca solve plus a forced solve
every 500 steps.  A step is one iteration of that loop for the tissue.

Prints one JSON line:
  value           cell-steps/s (sum over timed steps of the live cell count
                  divided by the wall time of the timed region)
  grid_cell_steps_per_s  active grid points swept (one ext_stim sweep per
                  step, 1000 ATP sweeps per Ca solve) per second
  roofline        the dominant kernel's algorithmic bytes per launch over its
                  mean duration (HIP events on the library's stream during
                  the timed region) against 8 TB/s
  cpu_baseline    the serial CPU restatement (oracle/) on a bounded sample
                  of the same workload, extrapolated to the timed region
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
# algorithmic bytes per unit of work: per grid point of a stencil sweep, and
# per AFM cell of a Ca sub-step 96 + 24 bytes per list entry (k_ca_cells)
BYTES_PER_POINT = {"k_atp": 22.0, "k_ext_stim": 21.0}
# per AFM cell: its own state read and written; per coupled list entry (a
# neighbour that is ALIVE, DEAD, FIX or MUSUME; the sub-step reads no other):
# the neighbour index 4, gj 8, the neighbour's (ca2p, IP3) 16; plus the gj
# write-back of an ALIVE neighbour, 8
CA_CELL_BYTES, CA_ENTRY_BYTES, CA_ALIVE_ENTRY_BYTES = 96.0, 28.0, 8.0
MODELED = ("k_ca_cells", "k_atp", "k_ext_stim", "k_ca_fused", "k_forces")
# A cell step moves 120 + 8 k bytes per cell (k = mean list length, the 6
# membrane slots included): positions read and written 48, per-cell
# attributes 40, bucket keys and permutation 16, renewal fields 16, the list
# written and read 8 k.  Split over the kernels that do each part:
STEP_GROUPS = (
    ("forces", ("k_forces", "k_forces_long"), "88 + 4 k per cell", lambda kb: 88.0 + 4.0 * kb),
    ("renewal", ("k_renew_flags", "k_renew_scan", "k_renew_apply", "k_remove", "k_pair_disperse",
                 "k_initialize_sc"), "16 per cell", lambda kb: 16.0),
    ("buckets", ("k_bucket_insert", "k_bucket_sort", "k_snapshot", "k_periodic_fix"),
     "16 per cell", lambda kb: 16.0),
    ("lists", ("k_sweep", "k_sweep_wave", "k_finish"), "4 k per cell", lambda kb: 4.0 * kb),
)
# the voxel maps, written once per step over the active box (cmap1 4 + cmap2
# 1 per voxel; the B stencil's 21 bytes per point count their reads), and
# the B stencil
MAP_KERNELS = ("k_zz_set", "k_zzmax", "k_map_clear", "k_map_stamp", "k_map_halo")
MAP_BYTES_PER_VOXEL = 5.0


def progress(msg):
    """a phase of the run, on stderr (the JSON line alone goes to stdout)"""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", choices=["large", "small", "batched", "adjoint-rd", "adjoint-mech", "ca-solve"],
                    default="large")
    ap.add_argument("--tissues", type=int, default=256, help="batched: number of tissues")
    ap.add_argument("--loop-sample", type=int, default=16,
                    help="batched: tissues timed one context at a time for the per-instance loop")
    ap.add_argument("--ca-every", type=int, default=500)
    ap.add_argument("--roofline-kernel", default="auto",
                    help="kernel whose roofline is reported (auto: the dominant one)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="approximate CPU time budget of the cpu_baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes instead of running the autotuner")
    ap.add_argument("--tuning", default=None,
                    help="JSON dict of launch shapes to use instead of running the autotuner "
                         "(profiling runs pin the shapes the timed run chose)")
    ap.add_argument("--copy-calibration", type=int, metavar="MIB", default=0,
                    help="only run a plain device copy of MIB MiB five times (the known-size kernel that "
                         "scales rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes)")
    ap.add_argument("--traffic-from", nargs=5, metavar=("KERNEL", "FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="summarise rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE output directories of this bench "
                         "and of --copy-calibration into bytes per launch of KERNEL (JSON on stdout)")
    ap.add_argument("--profile-all", action="store_true",
                    help="time every kernel with events and print a breakdown to stderr")
    ap.add_argument("--no-breakdown", action="store_true",
                    help="skip the untimed per-step kernel breakdown pass after the timed region")
    return ap.parse_args()


def window_start(warmup, steps, ca_every):
    """the driver's first step counter for a run of warmup untimed and
    `steps` timed steps whose last timed step runs a forced solve (forced
    solves follow the steps s with (s + 1) % ca_every == 0)"""
    return (-(warmup + steps)) % ca_every if ca_every > 0 else 0


def mean_list_length(c):
    """mean directed list length over all cells (the 6 membrane slots included)"""
    return float(np.diff(c["conn_ptr"]).mean()) if c["n"] else 0.0


def step_bytes(cells, kbar, points_b):
    """algorithmic bytes of one cell step: 120 + 8 k per cell, the maps' 5 per
    active-box voxel and the B stencil's 21 per point"""
    return cells * (120.0 + 8.0 * kbar) + (MAP_BYTES_PER_VOXEL + BYTES_PER_POINT["k_ext_stim"]) * points_b


def ca_solve_bytes(c, points_per_substep):
    """algorithmic bytes of one calc_ca2p solve (1000 sub-steps) of a tissue
    whose cells are c"""
    from epi_amd import configs as cf
    st, lens = c["state"], np.diff(c["conn_ptr"])
    afm = (st == cf.ALIVE) | (st == cf.FIX) | (st == cf.MUSUME)
    afm[:c["nmemb"] + c["nder"]] = False
    owner = np.repeat(np.arange(st.size), lens)
    nbs = st[c["conn"][:owner.size]]
    coupled = afm[owner] & np.isin(nbs, (cf.ALIVE, cf.DEAD, cf.FIX, cf.MUSUME))
    alive = coupled & (nbs == cf.ALIVE)
    per_sub = (CA_CELL_BYTES * int(afm.sum()) + CA_ENTRY_BYTES * int(coupled.sum())
               + CA_ALIVE_ENTRY_BYTES * int(alive.sum()) + BYTES_PER_POINT["k_atp"] * points_per_substep)
    return 1000.0 * per_sub


def step_breakdown(ts, native, drivers, steps):
    """An untimed pass of `steps` steps after the timed region with every
    kernel timed: per group of per-step kernels, its time per step, its
    algorithmic bytes per step and the bandwidth they imply."""
    t = 0
    ts.synchronize()
    a = ts.status(t)[0]
    ts.profile_reset()
    ts.profile_sample(1)
    ts.profile_enable(None)
    ts.step(drivers, steps)
    ts.synchronize()
    prof = ts.profile()
    ts.profile_enable([])
    ts.profile_reset()
    b = ts.status(t)[0]
    c = ts.get_cells(t)
    kbar = mean_list_length(c)
    n_steps = b.steps - a.steps
    cells = (b.cell_steps - a.cell_steps) / max(n_steps, 1)
    points_b = (b.grid_points_b - a.grid_points_b) / max(n_steps, 1)
    out = {}

    def add(name, kernels, model, byts):
        ms = sum(prof.get(k, (0.0, 0))[0] for k in kernels) / max(n_steps, 1)
        if ms <= 0:
            return
        out[name] = {"kernels": [k for k in kernels if k in prof], "us_per_step": round(1e3 * ms, 3),
                     "bytes_per_step": round(byts, 1), "model": model,
                     "achieved_gbs": round(byts / (ms * 1e-3) / 1e9, 2),
                     "frac": round(byts / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    for name, kernels, model, f in STEP_GROUPS:
        add(name, kernels, model, cells * f(kbar))
    add("maps", MAP_KERNELS, "5 per active-box voxel", MAP_BYTES_PER_VOXEL * points_b)
    add("ext_stim", ("k_ext_stim",), "21 per active point", BYTES_PER_POINT["k_ext_stim"] * points_b)
    tot_ms = sum(v[0] for k, v in prof.items()) / max(n_steps, 1)
    return {"steps": n_steps, "cells": round(cells, 1), "mean_list_length": round(kbar, 3),
            "active_b_points": round(points_b, 1), "kernel_us_per_step": round(1e3 * tot_ms, 3),
            "step_bytes": round(step_bytes(cells, kbar, points_b), 1), "groups": out}


# Events in the timed region: runs of SAMPLE_RUN consecutive launches, one
# run every SAMPLE_PERIOD launches.  A dispatch-recorded event pair brackets
# its kernel only when the launch before it carries events too, so the first
# launch of a run is not counted; sampling keeps the timer's cost out of the
# throughput.
SAMPLE_PERIOD, SAMPLE_RUN = 100, 5


def timed_region_profile(ts, kernels=None, sample=True):
    """dispatch-recorded events on `kernels` (None: every kernel) in the
    timed region, sampled"""
    ts.profile_reset()
    ts.profile_enable(kernels)
    ts.profile_sample(SAMPLE_PERIOD, SAMPLE_RUN) if sample else ts.profile_sample(1)


def dominant_kernel(prof, launched):
    """the modeled kernel with the largest estimated time in the timed
    region: its sampled mean duration times its launches there"""
    est = {k: prof[k][0] / prof[k][1] * launched.get(k, 0) for k in prof if prof[k][1] > 0}
    order = sorted(est, key=lambda k: -est[k])
    mix = {k: {"avg_us": round(1e3 * prof[k][0] / prof[k][1], 3), "launches": launched.get(k, 0),
               "timed_launches": prof[k][1], "est_ms": round(est[k], 4)} for k in order[:10]}
    return next((k for k in order if k in MODELED), None), mix


def roofline(ts, prof, s0, s1, rk, launched, workload):
    """Algorithmic bytes per launch of kernel rk over its mean launch
    duration (dispatch-recorded HIP events on the library's stream, sampled
    over the timed region).  `launched` counts every launch of rk in the
    timed region, timed or not."""
    if rk not in prof or not launched:
        return None
    ms, timed = prof[rk]
    avg_s = ms / timed / 1e3
    cnt = launched
    if rk in ("k_ca_cells", "k_ca_fused"):
        # a launch advances the cells of one Ca sub-step of every solving
        # tissue (a fused launch also the ATP grid of the sub-step before):
        # weight each tissue by its share of the launches
        per_solve = 1000 if rk == "k_ca_cells" else 999
        from epi_amd import configs as cf
        alg, units, entries, alive, points = 0.0, 0, 0.0, 0.0, 0.0
        for t in range(ts.T):
            solves = s1[t].ca_solves - s0[t].ca_solves
            if not solves:
                continue
            c = ts.get_cells(t)
            st, lens = c["state"], np.diff(c["conn_ptr"])
            afm = (st == cf.ALIVE) | (st == cf.FIX) | (st == cf.MUSUME)
            afm[:c["nmemb"] + c["nder"]] = False
            owner = np.repeat(np.arange(st.size), lens)
            nbs = st[c["conn"][:owner.size]]
            sel = afm[owner]
            coupled = sel & np.isin(nbs, (cf.ALIVE, cf.DEAD, cf.FIX, cf.MUSUME))
            w = solves * per_solve / cnt
            units += int(afm.sum() * w)
            entries += float(coupled.sum()) * w
            alive += float((coupled & (nbs == cf.ALIVE)).sum()) * w
            points += (s1[t].grid_points_ca - s0[t].grid_points_ca) / (solves * 1000) * w
        alg = CA_CELL_BYTES * units + CA_ENTRY_BYTES * entries + CA_ALIVE_ENTRY_BYTES * alive
        unit = (f"{units} AFM cells, {entries / max(units, 1):.2f} coupled list entries per cell "
                f"({alive / max(units, 1):.2f} ALIVE)")
        if rk == "k_ca_fused":
            alg += BYTES_PER_POINT["k_atp"] * points
            unit += f", {points:.0f} ATP grid points"
    elif rk == "k_forces":
        # one launch per step per context: every tissue's cells, 88 + 4 k each
        alg, cells_tot = 0.0, 0.0
        for t in range(ts.T):
            c = ts.get_cells(t)
            steps = max(s1[t].steps - s0[t].steps, 1)
            cells = (s1[t].cell_steps - s0[t].cell_steps) / steps
            cells_tot += cells
            alg += cells * STEP_GROUPS[0][3](mean_list_length(c))
        unit = f"{cells_tot:.0f} cells (88 + 4 k bytes each)"
    else:
        key = "grid_points_ca" if rk == "k_atp" else "grid_points_b"
        units = sum(getattr(b, key) - getattr(a, key) for a, b in zip(s0, s1)) / cnt
        alg = BYTES_PER_POINT[rk] * units
        unit = f"{units:.0f} grid points"
    achieved = alg / avg_s / 1e9
    traffic = None
    pmc = os.path.join(HERE, "profiles", f"{rk}_{workload}_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "kernel": rk,
            "avg_us": round(avg_s * 1e6, 3), "launches": cnt, "timed_launches": timed, "units_per_launch": unit,
            "algorithmic_bytes_per_launch": round(alg, 1)}
    return roof


def stencil_alone(ts, reps=20):
    """The B stencil (k_ext_stim) over the whole context on its own, after the
    timed region: in a step it runs on the side stream beside the next step's
    forces, which share the GPU with it, so its event span there is not its
    own speed.  Algorithmic bytes (21 per point below iz_bound) over the mean
    launch duration."""
    zz = ts.calc_zzmax()
    ts.calc_ext_stim(zz)
    ts.synchronize()
    a = [ts.status(t)[0] for t in range(ts.T)]
    ts.profile_reset()
    ts.profile_sample(1)
    ts.profile_enable(["k_ext_stim"])
    for _ in range(reps):
        ts.calc_ext_stim(zz)
    ts.synchronize()
    prof = ts.profile()
    ts.profile_enable([])
    b = [ts.status(t)[0] for t in range(ts.T)]
    if "k_ext_stim" not in prof:
        return None
    ms, cnt = prof["k_ext_stim"]
    avg_s = ms / cnt / 1e3
    pts = sum(y.grid_points_b - x.grid_points_b for x, y in zip(a, b)) / reps
    alg = BYTES_PER_POINT["k_ext_stim"] * pts
    traffic = None
    pmc = os.path.join(HERE, "profiles", "k_ext_stim_batched_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("bytes_per_launch")
    achieved = alg / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "kernel": "k_ext_stim",
            "avg_us": round(avg_s * 1e6, 3), "launches": reps, "units_per_launch": f"{pts:.0f} grid points",
            "algorithmic_bytes_per_launch": round(alg, 1)}


def make_spec(name):
    from epi_amd import configs
    return configs.config_large(2) if name == "large" else configs.config_small(1)


def n_draws_needed(spec, steps):
    st = spec["cells"]["state"]
    ready = int(np.count_nonzero((st == 4) | (st == 7)))
    return int((steps + 64) * (ready + 32) + 100000)


def cpu_threads():
    """threads the multithreaded CPU restatement runs on: OMP_NUM_THREADS if
    set, else every core this process may use"""
    n = os.environ.get("OMP_NUM_THREADS")
    if n and n.isdigit() and int(n) > 0:
        return int(n)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_rates(O, spec, seed, mt, step_runs, steps_per_run, sub_runs, subs_per_run, warm=10):
    """median seconds per cell step (sw gating held off, so no solve runs
    inside the sample) and per Ca sub-step of the CPU restatement on the
    workload's tissue, after `warm` excluded steps"""
    import statistics
    t = O.OracleTissue(spec["geometry"], mt=mt)
    t.load(spec)
    t.set_rng(O.mt19937_64_uniform(seed, n_draws_needed(spec, warm + step_runs * steps_per_run + 8)))
    drv = O.Driver(0, 0, 0, 0, 0)

    def one():
        t.step(drv)
        t.ca_calc_condition_reset()
    for _ in range(warm):
        one()
    per_step = []
    for _ in range(step_runs):
        t0 = time.perf_counter()
        for _ in range(steps_per_run):
            one()
        per_step.append((time.perf_counter() - t0) / steps_per_run)
    zz = t.calc_zzmax()
    per_sub = []
    for _ in range(sub_runs):
        t0 = time.perf_counter()
        t.calc_ca2p_n(zz, subs_per_run)
        per_sub.append((time.perf_counter() - t0) / subs_per_run)
    t.check()
    return statistics.median(per_step), statistics.median(per_sub)


def cpu_baseline(spec, seed, gpu_cell_steps, gpu_ca_solves, timed_steps, budget_s):
    """The CPU restatement (oracle/) on the host cores: the multithreaded
    build (OpenMP over the loops epi5_MT runs under TBB, CAS-added forces) on
    every thread, median of 3 runs of cell steps and of Ca sub-steps after 10
    excluded steps, extrapolated to the timed region's steps and solves; the
    serial build beside it on a shorter sample."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    threads = cpu_threads()
    # size the runs from one probe step and sub-step so that the whole sample
    # takes about budget_s
    probe_step, probe_sub = _cpu_rates(O, spec, seed, True, 1, 2, 1, 5, warm=1)
    k = max(5, min(400, int(budget_s * 0.4 / 3 / max(probe_step, 1e-6))))
    n_sub = max(10, min(1000, int(budget_s * 0.3 / 3 / max(probe_sub, 1e-6))))
    t_step, t_sub = _cpu_rates(O, spec, seed, True, 3, k, 3, n_sub)
    k_s = max(3, min(200, int(budget_s * 0.15 / max(probe_step * threads, 1e-6))))
    n_s = max(5, min(500, int(budget_s * 0.1 / max(probe_sub * threads, 1e-6))))
    s_step, s_sub = _cpu_rates(O, spec, seed, False, 1, k_s, 1, n_s)
    est = timed_steps * t_step + gpu_ca_solves * 1000 * t_sub
    est_serial = timed_steps * s_step + gpu_ca_solves * 1000 * s_sub
    return {
        "value": gpu_cell_steps / est,
        "unit": "cell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"multithreaded CPU restatement (OpenMP, {threads} threads, the loops epi5_MT runs under TBB) "
                   f"on the same tissue: median of 3 runs of {k} cell steps ({t_step*1e3:.2f} ms/step, sw gating "
                   f"held off) and of 3 runs of {n_sub} Ca sub-steps ({t_sub*1e3:.3f} ms each), after 10 excluded "
                   f"steps; extrapolated to the timed {timed_steps} steps + {gpu_ca_solves} Ca solves of 1000 "
                   f"sub-steps"),
        "seconds_per_step": t_step,
        "seconds_per_ca_substep": t_sub,
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "serial": {"value": gpu_cell_steps / est_serial, "cores": 1, "seconds_per_step": s_step,
                   "seconds_per_ca_substep": s_sub,
                   "sample": f"serial build: {k_s} cell steps and {n_s} Ca sub-steps after 10 excluded steps"},
    }


def run_batched(args):
    """256 independent tissues (uneven sheets, cell counts and grids) in one
    context: every kernel launch covers all of them (blockIdx.y = tissue).
    Beside it, the plain per-instance loop: a sample of the same tissues,
    each in its own context, run one after another over the same steps."""
    from epi_amd import autotune, configs, native
    specs = configs.batched_configs(args.tissues)
    geos = [s["geometry"] for s in specs]
    total = args.steps + args.warmup + args.ca_every
    draws = [native.mt19937_64_uniform(s["seed"], n_draws_needed(s, total)) for s in specs]
    # as in the plain run, the timed window ends on a forced solve
    start = window_start(args.warmup, args.steps, args.ca_every)
    drivers = lambda k: [native.Driver(start, 0, 0, args.ca_every, 0) for _ in range(k)]

    def make(idx):
        ts = native.TissueSet([geos[i] for i in idx])
        for t, i in enumerate(idx):
            ts.load(t, specs[i])
            ts.set_rng(t, draws[i])
        return ts

    tuning, tune_s = None, 0.0
    if args.tuning:
        tuning = json.loads(args.tuning)
    elif not args.no_tune:
        progress(f"autotuning on {len(specs)} tissues")
        tt = make(range(len(specs)))
        tuning, rep = autotune.tune(tt, drivers(len(specs)), warm_steps=1, b_reps=5, ca_sub=10,
                                    log=sys.stderr)
        tune_s = rep["seconds"]
        tt.close()
    ts = make(range(len(specs)))
    if tuning:
        ts.set_tuning(tuning)
    drv = drivers(len(specs))
    if args.warmup:
        ts.step(drv, args.warmup)
    ts.synchronize()
    s0 = [ts.status(t)[0] for t in range(len(specs))]
    progress(f"timing {args.steps} batch steps")
    timed_region_profile(ts, None)
    t0 = time.perf_counter()
    ts.step(drv, args.steps)
    ts.synchronize()
    wall = time.perf_counter() - t0
    s1 = [ts.status(t)[0] for t in range(len(specs))]
    prof = ts.profile()
    launched = ts.profile_launches()
    ts.profile_enable([])
    ts.profile_sample(1)
    cell_steps = sum(b.cell_steps - a.cell_steps for a, b in zip(s0, s1))
    grid_pts = sum((b.grid_points_b - a.grid_points_b) + (b.grid_points_ca - a.grid_points_ca)
                   for a, b in zip(s0, s1))
    ca_solves = sum(b.ca_solves - a.ca_solves for a, b in zip(s0, s1))
    errs = [b.err for b in s1 if b.err]
    dominant, kernels = dominant_kernel(prof, launched)
    rk = dominant if args.roofline_kernel == "auto" else args.roofline_kernel
    roof = roofline(ts, prof, s0, s1, rk, launched.get(rk, 0), "batched")
    b_alone = stencil_alone(ts)
    ts.close()
    # per-instance loop over an evenly spaced sample of the same tissues
    idx = sorted(set(int(round(i * (len(specs) - 1) / max(args.loop_sample - 1, 1)))
                     for i in range(args.loop_sample)))
    loop_cells, loop_wall = 0, 0.0
    progress(f"per-instance loop over {len(idx)} tissues")
    for i in idx:
        one = make([i])
        if tuning:
            one.set_tuning(tuning)
        d = drivers(1)
        # the same steps as the batch's timed region: its warm-up comes first
        one.step(d, args.warmup)
        one.synchronize()
        a = one.status(0)[0]
        t0 = time.perf_counter()
        one.step(d, args.steps)
        one.synchronize()
        loop_wall += time.perf_counter() - t0
        loop_cells += one.status(0)[0].cell_steps - a.cell_steps
        one.close()
    loop_value = loop_cells / loop_wall if loop_wall > 0 else None
    value = cell_steps / wall
    out = {
        "metric": "cell-steps/s + grid-cell-steps/s (FP64, 1e-10 vs ref); batched & adjoint vs plain",
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batched: {len(specs)} tissues in one context (sheet side U[32,96], "
                               f"free cells U[500,8000] capped to the box, grid edge U[32,96], nz = 2 edge, "
                               f"seeds 3000+t), sw-gated calc_ca2p plus one forced every {args.ca_every} "
                               f"steps, the window ending on a forced one",
                   "cells_start": int(sum(s["cells"]["x"].size for s in specs)),
                   "grid_points": int(sum((g.nx + 1) * (g.ny + 1) * (g.nz + 1) for g in geos)),
                   "ca_solves": int(ca_solves), "errors": errs, "launch_shapes": tuning,
                   "autotune_seconds": round(tune_s, 2)},
        "grid_cell_steps_per_s": round(grid_pts / wall, 1),
        "roofline": roof,
        "b_stencil_alone": b_alone,
        "kernel_mix": kernels,
        "per_instance_loop": None if loop_value is None else {
            "value": round(loop_value, 1), "unit": "cell-steps/s", "tissues_timed": len(idx), "steps": args.steps,
            "sample": "each sampled tissue in its own context, one after another, "
                      "timed over the same steps and Ca cadence as the batch"},
        "batched_vs_loop": None if loop_value is None else round(value / loop_value, 2),
    }
    print(json.dumps(out), flush=True)
    return 0


def copy_calibration(mib):
    """Five device copies of mib MiB (FP64): FETCH_SIZE and WRITE_SIZE of one
    such dispatch should read mib MiB each."""
    import torch
    n = mib * 1024 * 1024 // 8
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x5", flush=True)
    return 0


def _pmc_per_dispatch(d, counter, match, largest_grid=False):
    """per-dispatch values of `counter` for kernels whose name contains
    `match`, in dispatch order; with largest_grid only the dispatches of the
    largest grid (a batch's launches, not a single tissue's)"""
    import csv
    import glob
    vals, grid = {}, {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter or match not in row.get("Kernel_Name", ""):
                    continue
                key = (f, int(row.get("Dispatch_Id", 0)))
                vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
                grid[key] = int(row.get("Grid_Size") or 0)
    if largest_grid and grid:
        top = max(grid.values())
        vals = {k: v for k, v in vals.items() if grid[k] == top}
    return [v for _, v in sorted(vals.items(), key=lambda kv: kv[0][1])]


def traffic_summary(kernel, fdir, wdir, cfdir, cwdir, copy_mib=512):
    """HBM bytes per launch of `kernel` from the FETCH_SIZE / WRITE_SIZE runs
    (KiB per dispatch, over the kernel's dispatches of the largest grid), each
    scaled by what the counter reads for a copy of copy_mib MiB (the largest
    dispatches of the calibration run)."""
    import statistics
    copy_bytes = copy_mib * 1024 * 1024
    cf = statistics.median(sorted(_pmc_per_dispatch(cfdir, "FETCH_SIZE", ""))[-5:]) * 1024
    cw = statistics.median(sorted(_pmc_per_dispatch(cwdir, "WRITE_SIZE", ""))[-5:]) * 1024
    fetch = _pmc_per_dispatch(fdir, "FETCH_SIZE", kernel, largest_grid=True)
    write = _pmc_per_dispatch(wdir, "WRITE_SIZE", kernel, largest_grid=True)
    n = min(len(fetch), len(write), 1000)
    fk, wk = statistics.mean(fetch[-n:]), statistics.mean(write[-n:])
    return {"kernel": kernel, "dispatches": n, "fetch_kib_per_dispatch": fk, "write_kib_per_dispatch": wk,
            "copy_fetch_factor": copy_bytes / cf, "copy_write_factor": copy_bytes / cw,
            "bytes_per_launch": fk * 1024 * copy_bytes / cf + wk * 1024 * copy_bytes / cw}


def _timed(ts, fn):
    ts.synchronize()
    t0 = time.perf_counter()
    out = fn()
    ts.synchronize()
    return time.perf_counter() - t0, out


def run_ca_solve(args):
    """Whole calc_ca2p solves (1000 sub-steps) on the large tissue after a few
    steps: wall time per sub-step beside the kernels' event times."""
    from epi_amd import native
    spec = make_spec("large")
    ts = native.TissueSet([spec["geometry"]])
    ts.load(0, spec)
    ts.set_rng(0, native.mt19937_64_uniform(spec["seed"], n_draws_needed(spec, 64)))
    if args.tuning:
        ts.set_tuning(json.loads(args.tuning))
    elif not args.no_tune:
        from epi_amd import autotune
        tuning, _ = autotune.tune(ts, [native.Driver(0, 0, 0, 0, 0)], log=sys.stderr)
    ts.step([native.Driver(0, 0, 0, 0, 0)], 5)
    zz = ts.calc_zzmax()
    ts.calc_ca2p(zz)
    solves = 3
    t_wall, _ = _timed(ts, lambda: [ts.calc_ca2p(zz) for _ in range(solves)])
    ts.profile_reset()
    ts.profile_sample(1)
    ts.profile_enable(["k_ca_cells", "k_atp", "k_ca_fused"])
    ts.calc_ca2p(zz)
    ts.synchronize()
    prof = ts.profile()
    launched = ts.profile_launches()
    ts.profile_enable([])
    ts.profile_sample(1)
    out = {"metric": "calc_ca2p sub-step wall time", "value": round(1e6 * t_wall / (solves * 1000), 3), "unit": "us",
           "higher_is_better": False, "launch_shapes": ts.get_tuning(),
           "kernel_us": {k: round(1e3 * v[0] / v[1], 3) for k, v in prof.items()}}
    print(json.dumps(out), flush=True)
    ts.close()
    return 0


def window_mix(ts, fn):
    """Per-kernel time of one untimed call of fn with every kernel timed
    (dispatch-recorded events): {kernel: {"ms", "launches"}}."""
    ts.synchronize()
    ts.profile_reset()
    ts.profile_sample(1)
    ts.profile_enable(None)
    fn()
    ts.synchronize()
    prof = ts.profile()
    launched = ts.profile_launches()
    ts.profile_enable([])
    ts.profile_sample(1)
    ts.profile_reset()
    return {k: {"ms": round(ms, 4), "launches": cnt}
            for k, (ms, cnt) in sorted(prof.items(), key=lambda kv: -kv[1][0])}


def run_adjoint_rd(args):
    """Reaction-diffusion window of 200 calc_ca2p sub-steps on a 64x64x128
    grid with 5,000 frozen free cells; loss = mean squared AFM Ca2+ against a
    target from parameters perturbed by 5%; checkpoint every 20 sub-steps."""
    from epi_amd import configs, native
    spec = configs.config_adjoint_rd(4)
    g = spec["geometry"]
    ts = native.TissueSet([g])
    ts.load(0, spec)
    ts.set_rng(0, native.mt19937_64_uniform(4, 400000))
    ts.step([native.Driver(0, 0, 0, 0, 0)], 2)
    zz = ts.calc_zzmax()
    c = ts.get_cells(0)
    st, agek = c["state"], c["agek"]
    alive = st == 0
    kpa = np.where(alive, 4.0 + (6.0 - 4.0) * 0.5 * (1.0 + np.tanh((3.0 - agek) / 1.0)), 6.0)
    n_sub, every = 200, 20
    ts.rd_begin(zz)
    rng = np.random.default_rng(4)
    pert = dict(du=0.01 * 1.05, dp=0.1 * 0.95, gj0=0.99 * 0.95, kpa=kpa * (1 + 0.05 * rng.choice([-1, 1], kpa.size)))
    ts.rd_forward([pert], n_sub)
    target = ts.rd_get_ca2p(0)
    base = dict(du=0.01, dp=0.1, gj0=0.99, kpa=kpa)
    ts.rd_forward([base], n_sub)            # warm
    reps = 3
    t_fwd = min(_timed(ts, lambda: ts.rd_forward([base], n_sub))[0] for _ in range(reps))
    # lanes per cell of the reverse cell kernel: time both shapes, keep the faster
    shapes = {}
    for rg in (8, 16):
        ts.set_tuning({"rd_group": rg})
        ts.rd_adjoint([base], n_sub, every, [target])   # warm
        shapes[rg] = min((_timed(ts, lambda: ts.rd_adjoint([base], n_sub, every, [target]))
                          for _ in range(reps)), key=lambda r: r[0])
    rg_best = min(shapes, key=lambda k: shapes[k][0])
    ts.set_tuning({"rd_group": rg_best})
    t_adj, (loss, grads) = shapes[rg_best]
    mix = {"forward": window_mix(ts, lambda: ts.rd_forward([base], n_sub)),
           "adjoint": window_mix(ts, lambda: ts.rd_adjoint([base], n_sub, every, [target]))}
    ts.rd_end()
    gk = grads[0]["kpa"]
    out = {"metric": "adjoint cost / forward cost (reaction-diffusion window)", "value": round(t_adj / t_fwd, 3),
           "unit": "x", "higher_is_better": False, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "64x64 sheet + 16 FIX + 5,000 free cells frozen, grid 64x64x128, 200 Ca sub-steps, "
                                  "checkpoint every 20, seed 4", "cells": int(st.size), "iz_bound": int((zz[0] + 2.8) / (g.lz / g.nz))},
           "forward_ms": round(1e3 * t_fwd, 3), "adjoint_ms": round(1e3 * t_adj, 3),
           "adjoint_ms_by_rd_group": {str(k): round(1e3 * v[0], 3) for k, v in shapes.items()},
           "forward_substeps_per_s": round(n_sub / t_fwd, 1), "loss": float(loss[0]),
           "grad": {"du": grads[0]["du"], "dp": grads[0]["dp"], "gj0": grads[0]["gj0"],
                    "kpa_l2": float(np.linalg.norm(gk))},
           "kernel_mix": mix}
    print(json.dumps(out), flush=True)
    ts.close()
    return 0


def run_adjoint_mech(args):
    """Mechanics windows of 100 steps on 64 tissues of 2,000-6,000 cells
    (48^3 grids) in one context; loss = mean squared deviation of positions
    from a target made with parameters perturbed by 5%; checkpoint every 10."""
    from epi_amd import configs, native
    specs = configs.adjoint_mech_configs(64)
    ts = native.TissueSet([s["geometry"] for s in specs])
    for t, s in enumerate(specs):
        ts.load(t, s)
        ts.set_rng(t, native.mt19937_64_uniform(s["seed"], 300000))
    ts.step([native.Driver(0, 0, 0, 0, 0) for _ in specs], 2)
    n_steps, every = 100, 10
    ts.mech_begin()
    pert = [dict(eps=0.0105, k_adh=2.85, k_spring=26.25, k_bend=0.475) for _ in specs]
    ts.mech_forward(pert, n_steps)
    targets = [ts.mech_get_positions(t) for t in range(len(specs))]
    base = [dict() for _ in specs]
    ts.mech_forward(base, n_steps)   # warm
    reps = 3
    t_fwd = min(_timed(ts, lambda: ts.mech_forward(base, n_steps))[0] for _ in range(reps))
    ts.mech_adjoint(base, n_steps, every, targets)   # warm
    t_adj, (loss, grads) = min((_timed(ts, lambda: ts.mech_adjoint(base, n_steps, every, targets))
                                for _ in range(reps)), key=lambda r: r[0])
    mix = {"forward": window_mix(ts, lambda: ts.mech_forward(base, n_steps)),
           "adjoint": window_mix(ts, lambda: ts.mech_adjoint(base, n_steps, every, targets))}
    ts.mech_end()
    cells = sum(s["cells"]["x"].size for s in specs)
    out = {"metric": "adjoint cost / forward cost (batched mechanics windows)", "value": round(t_adj / t_fwd, 3),
           "unit": "x", "higher_is_better": False, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "64 tissues of 2,000-6,000 cells (40x40 sheet, grid 48^3), 100-step windows, "
                                  "checkpoint every 10, seeds 5000+t", "cells": int(cells)},
           "forward_ms": round(1e3 * t_fwd, 3), "adjoint_ms": round(1e3 * t_adj, 3),
           "forward_cell_steps_per_s": round(cells * n_steps / t_fwd, 1),
           "loss_mean": float(np.mean(loss)),
           "grad_mean": {k: float(np.mean([g[k] for g in grads])) for k in grads[0]},
           "kernel_mix": mix}
    print(json.dumps(out), flush=True)
    ts.close()
    return 0


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a batch of tissues would shard across GPUs by whole tissues, with no per-step exchange)",
              file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.traffic_from:
        print(json.dumps(traffic_summary(*args.traffic_from)))
        return 0
    if args.config == "batched":
        return run_batched(args)
    if args.config == "ca-solve":
        return run_ca_solve(args)
    if args.config == "adjoint-rd":
        return run_adjoint_rd(args)
    if args.config == "adjoint-mech":
        return run_adjoint_mech(args)
    return run_plain(args)


def run_plain(args):
    """One tissue stepped through proc.cpp:153-172.  The timed region is
    exactly args.steps steps and ends on a step whose forced calc_ca2p solve
    runs in it (the driver's step counter starts so that the cadence lands
    there): a short run times a solve in proportion one per window, a run of
    k * ca_every steps has k forced solves, as from step 0."""
    from epi_amd import native
    spec = make_spec(args.config)
    seed = spec["seed"]
    geo = spec["geometry"]
    total = args.steps + args.warmup + args.ca_every + 64
    draws = native.mt19937_64_uniform(seed, n_draws_needed(spec, total))
    tuning, tune_s = None, 0.0
    if args.tuning:
        tuning = json.loads(args.tuning)
    elif not args.no_tune:
        # launch shapes are chosen on a throwaway copy of the workload
        from epi_amd import autotune
        tt = native.TissueSet([geo])
        tt.load(0, spec)
        tt.set_rng(0, draws)
        progress("autotuning the launch shapes")
        tuning, rep = autotune.tune(tt, [native.Driver(0, 0, 0, 0, 0)], log=sys.stderr)
        tune_s = rep["seconds"]
        tt.close()
    ts = native.TissueSet([geo])
    ts.load(0, spec)
    ts.set_rng(0, draws)
    if tuning:
        ts.set_tuning(tuning)
    # forced solves fire after steps s with (s + 1) % ca_every == 0: start the
    # counter so that the last timed step is one of them
    start = window_start(args.warmup, args.steps, args.ca_every)
    drv = [native.Driver(start, 0, 0, args.ca_every, 0)]
    if args.warmup:
        ts.step(drv, args.warmup)
    ts.synchronize()
    s0, _ = ts.status(0)
    progress(f"timing {args.steps} steps")
    timed_region_profile(ts, None, sample=not args.profile_all)
    t0 = time.perf_counter()
    ts.step(drv, args.steps)
    ts.synchronize()
    t1 = time.perf_counter()
    s1, _ = ts.status(0)
    prof = ts.profile()
    launched = ts.profile_launches()
    ts.profile_enable([])
    ts.profile_sample(1)
    wall = t1 - t0
    cell_steps = s1.cell_steps - s0.cell_steps
    grid_pts = (s1.grid_points_b - s0.grid_points_b) + (s1.grid_points_ca - s0.grid_points_ca)
    ca_solves = s1.ca_solves - s0.ca_solves
    value = cell_steps / wall
    dominant, mix = dominant_kernel(prof, launched)
    rk = dominant if args.roofline_kernel == "auto" else args.roofline_kernel
    roof = roofline(ts, prof, [s0], [s1], rk, launched.get(rk, 0), args.config)
    # the whole window against the HBM roofline: every step's 120 + 8 k bytes
    # per cell, maps and B stencil, and every solve's cell and ATP bytes
    c = ts.get_cells(0)
    kbar = mean_list_length(c)
    n_steps = s1.steps - s0.steps
    alg = n_steps * step_bytes(cell_steps / max(n_steps, 1), kbar, (s1.grid_points_b - s0.grid_points_b) / max(n_steps, 1))
    if ca_solves:
        alg += ca_solves * ca_solve_bytes(c, (s1.grid_points_ca - s0.grid_points_ca) / (ca_solves * 1000))
    window = {"algorithmic_bytes": round(alg, 1), "achieved_gbs": round(alg / wall / 1e9, 2),
              "frac": round(alg / wall / 1e9 / PEAK_HBM_GBS, 4), "mean_list_length": round(kbar, 3),
              "note": "all algorithmic bytes of the timed steps and solves over the wall time of the timed region"}
    breakdown = step_breakdown(ts, native, drv, min(args.steps, 50)) if not args.no_breakdown else None
    if args.profile_all:
        tot = sum(v[0] for v in prof.values())
        for k, (ms, cnt) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
            print(f"  {k:18s} {ms:10.2f} ms {cnt:8d} launches {1e3*ms/cnt:9.2f} us/launch "
                  f"{100*ms/max(tot,1e-9):5.1f}%", file=sys.stderr)
    cpu = None
    if not args.no_cpu:
        progress("CPU baseline sample")
        cpu = cpu_baseline(spec, seed, cell_steps, ca_solves, args.steps, args.cpu_seconds)
    out = {
        "metric": "cell-steps/s + grid-cell-steps/s (FP64, 1e-10 vs ref); batched & adjoint vs plain",
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 {args.config} tissue ({geo.nmx}x{geo.nmy} sheet, "
                               f"{geo.nx}x{geo.ny}x{geo.nz} grid, seed {seed}), division and removal on, "
                               f"sw-gated calc_ca2p plus one forced every {args.ca_every} steps, the window "
                               f"ending on a forced one",
                   "cells_start": int(spec["cells"]["x"].size), "cells_end": int(s1.n),
                   "ca_solves": int(ca_solves), "launch_shapes": ts.get_tuning(),
                   "autotune_seconds": round(tune_s, 2)},
        "grid_cell_steps_per_s": round(grid_pts / wall, 1),
        "roofline": roof,
        "window_roofline": window,
        "cpu_baseline": cpu,
        "kernel_mix": mix,
        "step_breakdown": breakdown,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    ts.close()
    return 0

if __name__ == "__main__":
    sys.exit(main())
