"""setup_map_lat (map.cpp:18-167) against a vectorised brute force: a voxel
belongs to every cell whose lattice box contains it and whose sphere
(radius r; 2r for cmap2 of ALIVE/FIX/MUSUME) covers its lattice point; the
serial loop leaves the highest such cell index in cmap1."""
import numpy as np
import pytest

from epi_amd import configs
from helpers import pdiff


def circ(a, b, n):
    d = np.abs(a - b) % n
    return np.minimum(d, n - d)


@pytest.mark.parametrize("seed", [1, 2])
def test_maps_match_brute_force(oracle_mod, seed):
    spec = configs.config_tiny(seed, n_free=150, side=16, grid=20)
    g = spec["geometry"]
    t = oracle_mod.OracleTissue(g)
    t.load(spec)
    t.setup_map_lat()
    f = t.get_fields()
    c = t.get_cells()
    NX, NY, NZ = g.nx, g.ny, g.nz
    dx, dy, dz = g.lx / NX, g.ly / NY, g.lz / NZ
    irx, iry, irz = int(1.4 * NX / g.lx), int(1.4 * NY / g.ly), int(1.4 * NZ / g.lz)
    lat = c["lat"].reshape(-1, 3)
    assert np.array_equal(lat[:, 0], ((c["x"] * (NX / g.lx)).astype(int) + NX) % NX)
    assert np.array_equal(lat[:, 2], ((c["z"] * (NZ / g.lz)).astype(int) + NZ) % NZ)
    J, K, L = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    want1 = np.full((NX, NY, NZ), -1)
    want2 = np.zeros((NX, NY, NZ), np.uint8)
    for i in range(c["n"]):
        afm = c["state"][i] in (0, 4, 7)
        m = 2 if afm else 1
        inbox = ((circ(J, lat[i, 0], NX) <= m * irx) & (circ(K, lat[i, 1], NY) <= m * iry)
                 & (np.abs(L - lat[i, 2]) <= m * irz) & (L >= 1) & (L <= NZ - 1))
        d2 = (pdiff(J * dx, c["x"][i], g.lx) ** 2 + pdiff(K * dy, c["y"][i], g.ly) ** 2
              + (L * dz - c["z"][i]) ** 2)
        r = c["radius"][i]
        want1[inbox & (d2 < r * r)] = i
        if afm:
            want2[inbox & (d2 < (2 * r) ** 2)] = 1
    assert np.array_equal(f["cmap1"][:NX, :NY, :NZ], want1)
    assert np.array_equal(f["cmap2"][:NX, :NY, :NZ], want2)
    # periodic planes and the untouched top plane
    assert np.array_equal(f["cmap1"][NX, :, :NZ], f["cmap1"][0, :, :NZ])
    assert np.array_equal(f["cmap1"][:NX, NY, :NZ], f["cmap1"][:NX, 0, :NZ])
    assert np.array_equal(f["cmap2"][NX, :, :NZ], f["cmap2"][0, :, :NZ])
    assert np.all(f["cmap1"][:, :, NZ] == -1) and np.all(f["cmap2"][:, :, NZ] == 0)
    assert np.all(f["cmap1"][:, :, 0] == -1)
