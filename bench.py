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
MODELED = ("k_ca_cells", "k_atp", "k_ext_stim", "k_ca_fused")


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
    return ap.parse_args()


def kernel_mix(ts, native, steps):
    """An untimed pass of `steps` steps (one Ca period, so exactly one solve)
    with every kernel timed: the share of each kernel and the dominant one
    (the kernel with the largest total time if it has a byte model, else the
    costliest modeled one).  The timed region then times only that kernel."""
    ts.profile_reset()
    ts.profile_sample(1)
    ts.profile_enable(None)
    ts.step([native.Driver(0, 0, 0, steps, 0) for _ in range(ts.T)], steps)
    ts.synchronize()
    prof = ts.profile()
    launched = ts.profile_launches()
    ts.profile_enable([])
    ts.profile_sample(1)
    ts.profile_reset()
    tot = sum(v[0] for v in prof.values()) or 1.0
    order = sorted(prof.items(), key=lambda kv: -kv[1][0])
    mix = {k: {"avg_us": round(1e3 * ms / cnt, 3), "launches": cnt, "share": round(ms / tot, 4)}
           for k, (ms, cnt) in order[:8]}
    top = order[0][0] if order else None
    dominant = top if top in MODELED else next((k for k, _ in order if k in MODELED), None)
    return mix, dominant


# Events in the timed region: runs of SAMPLE_RUN consecutive launches, one
# run every SAMPLE_PERIOD launches.  A dispatch-recorded event pair brackets
# its kernel only when the launch before it carries events too, so the first
# launch of a run is not counted; sampling keeps the timer's cost out of the
# throughput.
SAMPLE_PERIOD, SAMPLE_RUN = 100, 5


def timed_region_profile(ts, rk, sample=True):
    ts.profile_reset()
    ts.profile_enable([rk])
    ts.profile_sample(SAMPLE_PERIOD, SAMPLE_RUN) if sample else ts.profile_sample(1)


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


def make_spec(name):
    from epi_amd import configs
    return configs.config_large(2) if name == "large" else configs.config_small(1)


def n_draws_needed(spec, steps):
    st = spec["cells"]["state"]
    ready = int(np.count_nonzero((st == 4) | (st == 7)))
    return int((steps + 64) * (ready + 32) + 100000)


def cpu_baseline(spec, seed, gpu_cell_steps, gpu_ca_solves, timed_steps, budget_s):
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    t.set_rng(O.mt19937_64_uniform(seed, n_draws_needed(spec, 400)))
    drv = O.Driver(0, 0, 0, 0, 0)
    # mechanics + B sample
    t0 = time.perf_counter()
    t.step(drv)
    t.step(drv)
    per = (time.perf_counter() - t0) / 2
    k = max(3, min(300, int(0.5 * budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    ran = 0
    for _ in range(k):
        ran += t.step(drv)
    t_mech = time.perf_counter() - t0
    zz = t.calc_zzmax()
    # Ca sample: a fraction of one solve, scaled to 1000 sub-steps
    t0 = time.perf_counter()
    t.calc_ca2p_n(zz, 10)
    per_sub = (time.perf_counter() - t0) / 10
    n_sub = max(10, min(1000, int(0.5 * budget_s / max(per_sub, 1e-6))))
    t0 = time.perf_counter()
    t.calc_ca2p_n(zz, n_sub)
    t_sub = (time.perf_counter() - t0) / n_sub
    t_step = t_mech / k
    est = timed_steps * t_step + gpu_ca_solves * 1000 * t_sub
    return {
        "value": gpu_cell_steps / est,
        "unit": "cell-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"serial CPU restatement on the same tissue: {k} steps at {t_step*1e3:.2f} ms/step "
                   f"(Ca solves in sample: {ran}) and {n_sub} Ca sub-steps at {t_sub*1e3:.3f} ms; "
                   f"extrapolated to {timed_steps} steps + {gpu_ca_solves} Ca solves"),
        "seconds_per_step": t_step,
        "seconds_per_ca_substep": t_sub,
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
    drivers = lambda k: [native.Driver(0, 0, 0, args.ca_every, 0) for _ in range(k)]

    def make(idx):
        ts = native.TissueSet([geos[i] for i in idx])
        for t, i in enumerate(idx):
            ts.load(t, specs[i])
            ts.set_rng(t, draws[i])
        return ts

    tuning, tune_s = None, 0.0
    if not args.no_tune:
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
    kernels, dominant = kernel_mix(ts, native, args.ca_every)
    rk = dominant if args.roofline_kernel == "auto" else args.roofline_kernel
    ts.synchronize()
    s0 = [ts.status(t)[0] for t in range(len(specs))]
    timed_region_profile(ts, rk)
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
    roof = roofline(ts, prof, s0, s1, rk, launched.get(rk, 0), "batched")
    ts.close()
    # per-instance loop over an evenly spaced sample of the same tissues
    idx = sorted(set(int(round(i * (len(specs) - 1) / max(args.loop_sample - 1, 1)))
                     for i in range(args.loop_sample)))
    loop_cells, loop_wall = 0, 0.0
    for i in idx:
        one = make([i])
        if tuning:
            one.set_tuning(tuning)
        d = drivers(1)
        # the same steps as the batch's timed region: its warm-up and its
        # untimed Ca period come first
        one.step(d, args.warmup + args.ca_every)
        one.synchronize()
        a = one.status(0)[0]
        t0 = time.perf_counter()
        one.step(d, args.steps)
        one.synchronize()
        loop_wall += time.perf_counter() - t0
        loop_cells += one.status(0)[0].cell_steps - a.cell_steps
        one.close()
    loop_value = loop_cells / loop_wall
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
                               f"seeds 3000+t), Ca every {args.ca_every} + sw gating",
                   "cells_start": int(sum(s["cells"]["x"].size for s in specs)),
                   "grid_points": int(sum((g.nx + 1) * (g.ny + 1) * (g.nz + 1) for g in geos)),
                   "ca_solves": int(ca_solves), "errors": errs, "launch_shapes": tuning,
                   "autotune_seconds": round(tune_s, 2)},
        "grid_cell_steps_per_s": round(grid_pts / wall, 1),
        "roofline": roof,
        "kernel_mix": kernels,
        "per_instance_loop": {"value": round(loop_value, 1), "unit": "cell-steps/s",
                              "tissues_timed": len(idx), "steps": args.steps,
                              "sample": "each sampled tissue in its own context, one after another, "
                                        "timed over the same steps and Ca cadence as the batch"},
        "batched_vs_loop": round(value / loop_value, 2),
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


def _pmc_per_dispatch(d, counter, match):
    import csv
    import glob
    vals = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter or match not in row.get("Kernel_Name", ""):
                    continue
                key = (f, int(row.get("Dispatch_Id", 0)))
                vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
    return [v for _, v in sorted(vals.items(), key=lambda kv: kv[0][1])]


def traffic_summary(kernel, fdir, wdir, cfdir, cwdir, copy_mib=512):
    """HBM bytes per launch of `kernel` from the FETCH_SIZE / WRITE_SIZE runs
    (KiB per dispatch), each scaled by what the counter reads for a copy of
    copy_mib MiB (the largest dispatches of the calibration run)."""
    import statistics
    copy_bytes = copy_mib * 1024 * 1024
    cf = statistics.median(sorted(_pmc_per_dispatch(cfdir, "FETCH_SIZE", ""))[-5:]) * 1024
    cw = statistics.median(sorted(_pmc_per_dispatch(cwdir, "WRITE_SIZE", ""))[-5:]) * 1024
    fetch = _pmc_per_dispatch(fdir, "FETCH_SIZE", kernel)
    write = _pmc_per_dispatch(wdir, "WRITE_SIZE", kernel)
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
    if not args.no_tune:
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
    from epi_amd import native
    spec = make_spec(args.config)
    seed = spec["seed"]
    geo = spec["geometry"]
    draws = native.mt19937_64_uniform(seed, n_draws_needed(spec, args.steps + args.warmup + args.ca_every))
    tuning, tune_s = None, 0.0
    if args.tuning:
        tuning = json.loads(args.tuning)
    elif not args.no_tune:
        # launch shapes are chosen on a throwaway copy of the workload
        from epi_amd import autotune
        tt = native.TissueSet([geo])
        tt.load(0, spec)
        tt.set_rng(0, draws)
        tuning, rep = autotune.tune(tt, [native.Driver(0, 0, 0, 0, 0)], log=sys.stderr)
        tune_s = rep["seconds"]
        tt.close()
    ts = native.TissueSet([geo])
    ts.load(0, spec)
    ts.set_rng(0, draws)
    if tuning:
        ts.set_tuning(tuning)
    drv = [native.Driver(0, 0, 0, args.ca_every, 0)]
    if args.warmup:
        ts.step(drv, args.warmup)
    kernels, dominant = kernel_mix(ts, native, args.ca_every)
    rk = dominant if args.roofline_kernel == "auto" else args.roofline_kernel
    ts.synchronize()
    s0, _ = ts.status(0)
    if args.profile_all:
        ts.profile_reset()
        ts.profile_enable(None)
    else:
        timed_region_profile(ts, rk)
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
    roof = roofline(ts, prof, [s0], [s1], rk, launched.get(rk, 0), args.config)
    if args.profile_all:
        tot = sum(v[0] for v in prof.values())
        for k, (ms, cnt) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
            print(f"  {k:18s} {ms:10.2f} ms {cnt:8d} launches {1e3*ms/cnt:9.2f} us/launch "
                  f"{100*ms/max(tot,1e-9):5.1f}%", file=sys.stderr)
    cpu = None
    if not args.no_cpu:
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
                               f"{geo.nx}x{geo.ny}x{geo.nz} grid, seed {seed}), Ca every {args.ca_every} + sw gating",
                   "cells_start": int(spec["cells"]["x"].size), "cells_end": int(s1.n),
                   "ca_solves": int(ca_solves), "launch_shapes": ts.get_tuning(),
                   "autotune_seconds": round(tune_s, 2)},
        "grid_cell_steps_per_s": round(grid_pts / wall, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_mix": kernels,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    ts.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
