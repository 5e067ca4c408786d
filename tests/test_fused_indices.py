"""The index arithmetic of the fused Ca launch, restated on the host and
checked against the grid and LDS bounds on the parity workloads (CPU only).

k_ca_corners and the cell part of k_ca_fused read 16 grid corners per AFM
cell: corner k = q & 7 of grid_avg8 (utils.h:18-23) at the cell's lattice
point plus the offsets below, slots 0-7 from the B field and 8-15 from ATP.
Below iz_bound an ATP corner is recomputed from its seven-point stencil
(periodic x and y, mirror z, ca2p.cpp:302-338).  Every one of those reads
must fall inside the (NX+1)(NY+1)(NZ+1) grid of pitch XP, plane PL; the
per-lane chunk slots must fall inside the leader's LDS rows."""
import numpy as np
import pytest

from epi_amd import configs

# corner k -> (dx, dy, dz), as the kernels spell it
OFF = np.array([[int(k in (1, 4, 5, 7)), int(k in (2, 4, 6, 7)), int(k in (3, 5, 6, 7))] for k in range(8)])


def lattice(c, nx, ny, nz, lx, ly, lz):
    """set_lattice (map.cpp:18-22) as k_map_stamp computes it: truncation
    toward zero, then i % N for i in [-N, 2N)"""
    out = []
    for v, n, L in ((c["x"], nx, lx), (c["y"], ny, ly), (c["z"], nz, lz)):
        i = np.trunc(v * (n / L)).astype(np.int64)
        assert np.all((i >= -n) & (i < 2 * n))
        out.append(np.mod(i, n))
    return np.stack(out, 1)


def check_reads(spec, margin=2.8):
    g = spec["geometry"]
    c = spec["cells"]
    NX, NY, NZ = g.nx, g.ny, g.nz
    XP = (NX + 1 + 15) // 16 * 16
    PL = XP * (NY + 1)
    gsize = PL * (NZ + 1)
    start = spec["nmemb"] + spec["nder"]
    st = c["state"]
    afm = np.isin(st, (configs.ALIVE, configs.FIX, configs.MUSUME))
    afm[:start] = False
    lat = lattice(c, NX, NY, NZ, g.lx, g.ly, g.lz)[afm]
    izb = int((c["z"].max() + margin) * (NZ / g.lz))
    assert 0 < izb <= NZ
    reads = []
    for k in range(8):
        ix, iy, iz = (lat + OFF[k]).T
        # B corner and a plain ATP corner: the grid point itself, copies at NX, NY
        assert ix.max() <= NX and iy.max() <= NY and iz.max() <= NZ and min(ix.min(), iy.min(), iz.min()) >= 0
        reads.append(iz * PL + iy * XP + ix)
        # recomputed ATP corner: x = NX, y = NY fold to 0; its stencil below iz_bound
        j = np.where(ix == NX, 0, ix)
        kk = np.where(iy == NY, 0, iy)
        sel = iz < izb
        j, kk, l = j[sel], kk[sel], iz[sel]
        assert l.max(initial=0) <= NZ - 1
        px, nxx = np.where(j == 0, NX - 1, j - 1), np.where(j == NX - 1, 0, j + 1)
        py, nyy = np.where(kk == 0, NY - 1, kk - 1), np.where(kk == NY - 1, 0, kk + 1)
        pz, nzz = np.where(l == 0, 1, l - 1), np.where(l + 1 == NZ, NZ - 1, l + 1)
        for a, b, d in ((j, kk, l), (px, kk, l), (nxx, kk, l), (j, py, l), (j, nyy, l), (j, kk, pz), (j, kk, nzz)):
            assert a.min(initial=0) >= 0 and a.max(initial=0) <= NX - 1
            assert b.min(initial=0) >= 0 and b.max(initial=0) <= NY - 1
            assert d.min(initial=0) >= 0 and d.max(initial=0) <= NZ - 1
            reads.append(d * PL + b * XP + a)
    r = np.concatenate(reads)
    assert r.min() >= 0 and r.max() < gsize
    return int(afm.sum())


@pytest.mark.parametrize("make", [
    lambda: configs.config_tiny(7, 160),
    lambda: configs.config_tiny(21, n_free=300, side=16, grid=16, n_layers=3, dead_top=True, nz=40),
    lambda: configs.config_small(1),
    lambda: configs.config_large(2),
    lambda: configs.config_dense_sheet(11, 26),
])
def test_corner_reads_stay_in_the_grid(make):
    assert check_reads(make()) > 0


@pytest.mark.parametrize("cg", [4, 8, 16])
def test_chunk_slots_stay_in_the_leader_rows(cg):
    """EPL entries per lane (4 with 4 lanes, else 2), CCAP = EPL CG entries per
    chunk in rows of CCAP + 1 double2; NQ = 16 / CG corners per lane in rows
    of 17; the first chunk is read before the row length, so a row must hold
    CCAP slots (the row stride is at least 64)"""
    epl = 4 if cg < 8 else 2
    ccap = epl * cg
    slots = np.array([lane + i * cg for lane in range(cg) for i in range(epl)])
    assert slots.min() == 0 and slots.max() == ccap - 1 and np.unique(slots).size == ccap
    nq = 16 // cg
    corners = np.array([lane + j * cg for lane in range(cg) for j in range(nq)])
    assert np.array_equal(np.sort(corners), np.arange(16))
    assert ccap <= 64
