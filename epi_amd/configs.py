"""Seeded synthetic tissues for the epi5_MT hot path.

A tissue spec is a plain dict of numpy arrays in the reference's load order
(MEMB rows, then DER rows, then the rest), the same content the reference's
17-column input deck carries (cellmanager.cpp:243-302), plus resume-style
per-cell values (the load_uvp path, cell_init.cpp:47-71) and initial fields
(read_binary into both buffers, proc.cpp:107-110).

Generator decisions (applied identically to every backend):
  * sheet: ``nmx`` columns x ``nmy`` rows (nmy even, the triangular adjacency
    of cellmanager.cpp:66-113 needs it), x pitch ``lx/nmx``, y pitch
    ``ly/nmy``, odd rows shifted by half a pitch in x, centres at
    z = R_memb = 1.0 (a literal z = 0 divides by zero in wall_interaction,
    cell_interaction.cpp:287-294);
  * stem cells (FIX) on a regular sub-lattice of the first free layer at
    z = 1 + 1 + 1.4, ``rest_div_times`` 15 (cell.cpp:18-20), ageb ~ U[4.5, 6)
    so they are division-ready (cell_state_renew.cpp:147-155);
  * free cells on jittered (+-0.1 per axis) square lattices with an in-plane
    pitch of at least 2.0 and a layer pitch of 2.2; the number of layers is
    the requested one or, if that would put cells closer than 2.0, as many as
    needed (two layers of 1,000 cells over a 51.2 x 51.2 sheet means a 1.6
    pitch, where one explicit step moves overlapping radius-1.4 cells through
    each other and the reference aborts with |coef| > 1000,
    cell_interaction.cpp:51-55):
    layer 1 MUSUME (rest_div_times 3, ageb ~ U[4.5, 6)), upper layers ALIVE
    (agek ~ U[0, 22), a few in [21.99, 22) so cornification and the sw-gated
    Ca solve happen early), optionally a DEAD top layer (agek ~ U[31.1, 31.3),
    lifted by 0.7 so its contact count allows removal) placed before the last
    ALIVE layer in index order so removals never move a queued cell;
  * fields ATP and B = 0.1 + U(+-0.01) in both buffers; per-cell Ca2+ and IP3
    likewise; ex_inert 0.97; ca2p_avg = Ca2+;
  * replay RNG stream: std::mt19937_64(seed) through
    uniform_real_distribution<double>(0,1); generator randomness comes from
    numpy's PCG64 seeded with seed ^ 0x9E3779B97F4A7C15.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, asdict

import numpy as np

ALIVE, DEAD, DISA, UNUSED, FIX, BLANK, DER, MUSUME, AIR, MEMB = range(10)
STATE_NAMES = ["ALIVE", "DEAD", "DISA", "UNUSED", "FIX", "BLANK", "DER",
               "MUSUME", "AIR", "MEMB"]

R_MAX = 1.4
R_MEMB = 1.0
CELL_FIELDS = ("state", "fix_origin", "x", "y", "z", "radius", "ca2p",
               "ca2p_avg", "ip3", "ex_inert", "agek", "ageb", "ex_fat",
               "in_fat", "spr_nat_len", "rest_div_times", "is_malignant",
               "pair")


@dataclass
class Geometry:
    nmx: int
    nmy: int
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float
    max_cells: int

    def as_dict(self):
        return asdict(self)

    @property
    def grid_shape(self):
        return (self.nx + 1, self.ny + 1, self.nz + 1)

    @property
    def buckets(self):
        # cell_connection.cpp:22-26
        return (int(self.lx / 2.0 + 0.5), int(self.ly / 2.0 + 0.5),
                int(self.lz / 2.0))


def default_geometry(max_cells=None):
    """epi5_MT's compile-time geometry (define.h:72-113, NMY_tri = 114)."""
    return Geometry(nmx=100, nmy=114, nx=100, ny=100, nz=200, lx=50.0,
                    ly=50.0, lz=100.0,
                    max_cells=max_cells or 100 * 114 + 50000)


def check_bucket_span(lx):
    """True when every x in [0, lx) maps below ANX (the reference asserts
    otherwise, cell_connection.cpp:48-53)."""
    anx = int(lx / 2.0 + 0.5)
    top = math.nextafter(lx, 0.0)
    return int(top / (2.0 + 1e-7)) < anx


def _empty_cells(n):
    c = {k: np.zeros(n, dtype=np.float64) for k in CELL_FIELDS}
    for k in ("state", "is_malignant"):
        c[k] = np.zeros(n, dtype=np.uint8)
    for k in ("fix_origin", "rest_div_times", "pair"):
        c[k] = np.zeros(n, dtype=np.int32)
    c["pair"][:] = -1
    c["fix_origin"][:] = -1
    return c


def _concat(parts):
    return {k: np.concatenate([p[k] for p in parts]) for k in CELL_FIELDS}


def make_sheet(nmx, nmy, lx, ly, z=R_MEMB):
    assert nmy % 2 == 0
    px, py = lx / nmx, ly / nmy
    cols = np.tile(np.arange(nmx), nmy)
    rows = np.repeat(np.arange(nmy), nmx)
    c = _empty_cells(nmx * nmy)
    c["state"][:] = MEMB
    c["x"] = cols * px + np.where(rows % 2 == 1, 0.5 * px, 0.0)
    c["y"] = rows * py
    c["z"][:] = z
    c["radius"][:] = R_MEMB
    return c


def _lattice(rng, side_n, lx, ly, z, jitter=0.1):
    pitch_x, pitch_y = lx / side_n, ly / side_n
    a = np.tile(np.arange(side_n), side_n)
    b = np.repeat(np.arange(side_n), side_n)
    x = (a + 0.5) * pitch_x + rng.uniform(-jitter, jitter, a.size)
    y = (b + 0.5) * pitch_y + rng.uniform(-jitter, jitter, a.size)
    zz = z + rng.uniform(-jitter, jitter, a.size)
    return x, y, zz


def make_tissue(seed, nmx, nmy, lx, ly, nx, ny, nz, n_fix, n_free, n_layers,
                dead_top=False, n_hot_alive=30, max_cells=None, lz=None,
                min_pitch=2.0):
    """Build one synthetic tissue spec (see module docstring)."""
    rng = np.random.default_rng(seed ^ 0x9E3779B97F4A7C15)
    lz = lz if lz is not None else nz * (lx / nx)
    geo = Geometry(nmx=nmx, nmy=nmy, nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz,
                   max_cells=0)
    sheet = make_sheet(nmx, nmy, lx, ly)
    side_n = max(2, int(min(lx, ly) / min_pitch))
    sites = side_n * side_n
    n_layers = max(n_layers, int(math.ceil((n_free + n_fix) / sites)))
    per_layer = [n_free // n_layers] * n_layers
    per_layer[0] += n_free - sum(per_layer)
    # layer 0 shares its sites with the FIX cells; the excess moves up
    extra = max(0, per_layer[0] + n_fix - sites)
    per_layer[0] -= extra
    for li in range(1, n_layers):
        if not extra:
            break
        move = min(extra, sites - per_layer[li])
        per_layer[li] += move
        extra -= move
    assert extra == 0 and all(0 < p <= sites for p in per_layer), "tissue does not fit its layers"
    z0 = R_MEMB + R_MEMB + R_MAX
    layers = []
    fix_part = None
    for li in range(n_layers):
        z = z0 + 2.2 * li
        if dead_top and li == n_layers - 1:
            z += 0.7
        x, y, zz = _lattice(rng, side_n, lx, ly, z)
        idx = np.arange(x.size)
        if li == 0 and n_fix > 0:
            fs = int(round(math.sqrt(n_fix)))
            assert fs * fs == n_fix, "n_fix must be a square"
            step = side_n / fs
            fa = (np.arange(fs) * step + step / 2).astype(int)
            fix_sites = (fa[None, :] + side_n * fa[:, None]).ravel()
            fix_part = _empty_cells(n_fix)
            fix_part["state"][:] = FIX
            fix_part["x"], fix_part["y"] = x[fix_sites], y[fix_sites]
            fix_part["z"][:] = z0
            fix_part["radius"][:] = R_MAX
            fix_part["rest_div_times"][:] = 15
            fix_part["ageb"] = rng.uniform(4.5, 6.0, n_fix)
            fix_part["fix_origin"] = np.arange(n_fix, dtype=np.int32)
            idx = np.setdiff1d(idx, fix_sites)
        take = np.sort(rng.choice(idx, per_layer[li], replace=False))
        part = _empty_cells(per_layer[li])
        part["x"], part["y"], part["z"] = x[take], y[take], zz[take]
        part["radius"][:] = R_MAX
        part["fix_origin"] = rng.integers(0, max(n_fix, 1), per_layer[li]).astype(np.int32)
        if li == 0:
            part["state"][:] = MUSUME
            part["rest_div_times"][:] = 3
            part["ageb"] = rng.uniform(4.5, 6.0, per_layer[li])
        elif dead_top and li == n_layers - 1:
            part["state"][:] = DEAD
            part["agek"] = rng.uniform(31.1, 31.3, per_layer[li])
        else:
            part["state"][:] = ALIVE
            part["agek"] = rng.uniform(0.0, 22.0, per_layer[li])
        layers.append(part)
    alive_layers = [i for i, p in enumerate(layers) if p["state"][0] == ALIVE]
    if alive_layers and n_hot_alive:
        hot = layers[alive_layers[0]]
        k = min(n_hot_alive, hot["agek"].size)
        sel = rng.choice(hot["agek"].size, k, replace=False)
        hot["agek"][sel] = rng.uniform(21.99, 22.0, k)
    # index order: MEMB, FIX, MUSUME layer, ALIVE layers, DEAD layer, last ALIVE layer
    order = list(range(n_layers))
    if dead_top and n_layers >= 3:
        order = list(range(n_layers - 2)) + [n_layers - 1, n_layers - 2]
    parts = [sheet] + ([fix_part] if fix_part is not None else []) + [layers[i] for i in order]
    cells = _concat(parts)
    n = cells["x"].size
    cells["x"] = np.mod(cells["x"], lx)
    cells["y"] = np.mod(cells["y"], ly)
    geo.max_cells = max_cells or int(n * 1.5 + 1024)
    # resume-style values (load_uvp): Ca2+/IP3 noise, ex_inert, ca2p_avg
    ca = 0.1 + rng.uniform(-0.01, 0.01, n)
    ip3 = 0.1 + rng.uniform(-0.01, 0.01, n)
    cells["ca2p"] = ca.copy()
    cells["ca2p_avg"] = ca.copy()
    cells["ip3"] = ip3.copy()
    cells["ex_inert"][:] = 0.97
    shape = geo.grid_shape
    atp = 0.1 + rng.uniform(-0.01, 0.01, shape)
    b = 0.1 + rng.uniform(-0.01, 0.01, shape)
    for f in (atp, b):
        f[nx, :, :] = f[0, :, :]
        f[:, ny, :] = f[:, 0, :]
    return {
        "geometry": geo,
        "nmemb": nmx * nmy,
        "nder": 0,
        "cells": cells,
        "values": {"ca2p": ca, "ip3": ip3, "ex_inert": cells["ex_inert"].copy(),
                   "ca2p_avg": ca.copy()},
        "fields": {"atp": atp, "b": b},
        "seed": seed,
    }


def config_small(seed=1):
    """Plain FP64, small: periodic 64x64 sheet (pitch 0.8), 16 FIX,
    2,000 free cells (4 layers at pitch 2.0, see the module docstring),
    grid 64^3 (dx 0.8), seed 1."""
    return make_tissue(seed, nmx=64, nmy=64, lx=51.2, ly=51.2, nx=64, ny=64,
                       nz=64, n_fix=16, n_free=2000, n_layers=2)


def config_large(seed=2):
    """Plain FP64, large single tissue: 128x128 sheet, 64 FIX, 20,000 free
    cells in 8 layers (top layer DEAD), grid 128x128x256, seed 2.  The sheet
    pitch is 0.8125 (L = 104) rather than 0.8 (L = 102.4): with L = 102.4
    cells in the last 0.4 of the domain fall in bucket ANX and the reference
    aborts (cell_connection.cpp:22-26, 48-53)."""
    return make_tissue(seed, nmx=128, nmy=128, lx=104.0, ly=104.0, nx=128,
                       ny=128, nz=256, n_fix=64, n_free=20000, n_layers=8,
                       dead_top=True, n_hot_alive=60)


def config_dense_sheet(seed=11, sheet=26):
    """Small tissue on a dense membrane sheet: sheet x sheet particles on a
    14x14 box (pitch 0.54 for 26), 300 free cells in 3 layers, grid
    16x16x32.  A basal cell's list then holds the many sheet particles
    within its reach, without the strong interactions that make the
    reference abort.  With sheet 26 the longest list is 46 at the first
    build and 60 after 120 steps; with 28, 50 and 66."""
    lx = 14.0
    return make_tissue(seed, nmx=sheet, nmy=sheet, lx=lx, ly=lx, nx=16, ny=16, nz=32,
                       n_fix=4, n_free=300, n_layers=3, n_hot_alive=4, lz=32 * lx / 16)


def config_tiny(seed=7, n_free=160, side=16, grid=16, n_layers=2, dead_top=False, nz=None):
    """Small tissue for fast parity tests (same generator); nz > grid makes
    the box taller (dz = dx)."""
    lx = 0.8 * side
    if not check_bucket_span(lx):
        lx = 2.0 * math.ceil(0.4 * side)
    nz = nz or grid
    return make_tissue(seed, nmx=side, nmy=side, lx=lx, ly=lx, nx=grid,
                       ny=grid, nz=nz, n_fix=4, n_free=n_free,
                       n_layers=n_layers, dead_top=dead_top, n_hot_alive=4,
                       lz=nz * (lx / grid))


def batched_capacity(side_n, lz, dz, n_fix):
    """Free cells a box holds at the layer pitch of make_tissue while the top
    layer keeps calc_ca2p's iz_bound = (zzmax + 2.8) / dz below nz, with a
    margin of 6.4: the jitter, plus two layer pitches for the tissue's upward
    drift (with a margin of 2.0, the 32 x 32 x 64 box of tissue 91 left its
    grid within the batched benchmark's 700-odd steps)."""
    z0 = R_MEMB + R_MEMB + R_MAX
    layers = int((lz - 2.8 - 2.0 * dz - 6.4 - z0) / 2.2) + 1
    return side_n * side_n * max(layers, 1) - n_fix


def batched_configs(n_tissues=256, base_seed=3000, scale=1.0):
    """Batched config: per tissue t (seed 3000+t) a sheet side ~U[32,96]
    (even), free cells ~U[500,8000], 16 FIX, grid edge e ~U[32,96],
    nz = 2e, dx = dy = dz = L/e (so LZ = 2L).  L = 2*ceil(0.4*side) keeps
    every x below ANX.  The free-cell count is capped at what the box holds
    at layer pitch 2.0 x 2.2 (batched_capacity): a side-32 sheet (L = 26)
    holds about 3,000.  ``scale`` < 1 shrinks the cell counts for quick
    tests."""
    out = []
    for t in range(n_tissues):
        r = np.random.default_rng(base_seed + t)
        side = int(r.integers(16, 49)) * 2
        n_free = int(r.integers(500, 8001) * scale)
        edge = int(r.integers(32, 97))
        lx = 2.0 * math.ceil(0.4 * side)
        lz = 2 * edge * (lx / edge)
        side_n = max(2, int(lx / 2.0))
        n_free = max(64, min(n_free, batched_capacity(side_n, lz, lx / edge, 16)))
        out.append(make_tissue(base_seed + t, nmx=side, nmy=side, lx=lx, ly=lx,
                               nx=edge, ny=edge, nz=2 * edge, n_fix=16,
                               n_free=n_free, n_layers=2, n_hot_alive=8, lz=lz))
    return out


def config_adjoint_rd(seed=4):
    """Adjoint reaction-diffusion workload: a 64x64 sheet (pitch 0.8), 16 FIX
    and 5,000 free cells (positions frozen during the window), FP64 grid
    64x64x128 (dx = dz = 0.8), seed 4."""
    return make_tissue(seed, nmx=64, nmy=64, lx=51.2, ly=51.2, nx=64, ny=64, nz=128,
                       n_fix=16, n_free=5000, n_layers=3, n_hot_alive=0, lz=102.4)


def adjoint_mech_configs(n_tissues=64, base_seed=5000):
    """Batched mechanics adjoint workload: per tissue t (seed 5000+t) a
    40x40 sheet (L = 32), 16 FIX and free cells bringing the tissue to
    U[2000, 6000] cells, FP64 grid 48^3 (dx = L/48, dz = LZ/48 with
    LZ = 64 so that the stack fits the box)."""
    out = []
    for t in range(n_tissues):
        r = np.random.default_rng(base_seed + t)
        n_total = int(r.integers(2000, 6001))
        n_free = n_total - 40 * 40 - 16
        out.append(make_tissue(base_seed + t, nmx=40, nmy=40, lx=32.0, ly=32.0, nx=48, ny=48, nz=48,
                               n_fix=16, n_free=n_free, n_layers=2, n_hot_alive=0, lz=64.0))
    return out
