"""The 7-point FP64 stencils: ext_stim B (ext_stim.cpp:18-69) and ATP
(ca2p.cpp:302-338), against closed forms on an anisotropic grid
(dx = 1, dz = 0.5): B weighs every neighbour by inv_dz^2, ATP by inv_dx^2."""
import numpy as np
import pytest

from helpers import sheet_spec

DT, DB, KB, DA, KAA = 0.01, 0.0009, 0.025, 1.0, 0.5


def grid_tissue(O, NX=16, NZ=8, LZ=4.0):
    spec = sheet_spec([], nmx=4, nmy=4, L=float(NX), grid=(NX, NX, NZ), lz=LZ)
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    shape = (NX + 1, NX + 1, NZ + 1)
    t.set_maps(cmap1=np.full(shape, -1, np.int32), cmap2=np.ones(shape, np.uint8))
    return t, spec["geometry"], shape


def zz_for(g, izb, ca=False):
    dz = g.lz / g.nz
    zz = izb * dz - 2.8 + 0.3 * dz
    assert (int((zz + 2.8) / dz) if ca else int((zz + 2.8) * (g.nz / g.lz))) == izb
    return zz


@pytest.mark.parametrize("k", [1, 3])
def test_b_fourier_mode_uses_inv_dz(oracle_mod, k):
    t, g, shape = grid_tissue(oracle_mod)
    j = np.arange(shape[0])
    B = np.broadcast_to(np.cos(2 * np.pi * k * (j % g.nx) / g.nx)[:, None, None], shape).copy()
    t.set_fields(b=B, b_next=B)
    t.calc_ext_stim(zz_for(g, g.nz))
    out = t.get_fields()["b"]
    inv_dz2 = (g.nz / g.lz) ** 2
    lam = 1 + DT * (DB * inv_dz2 * (2 * np.cos(2 * np.pi * k / g.nx) - 2) - KB)
    np.testing.assert_allclose(out[:g.nx, :g.ny, :g.nz], (lam * B)[:g.nx, :g.ny, :g.nz], rtol=0, atol=1e-14)
    # periodic copies of planes x = NX and y = NY
    assert np.array_equal(out[g.nx, :, :g.nz], out[0, :, :g.nz])
    assert np.array_equal(out[:g.nx, g.ny, :g.nz], out[:g.nx, 0, :g.nz])


def test_b_z_boundaries(oracle_mod):
    t, g, shape = grid_tissue(oracle_mod)
    l = np.arange(shape[2], dtype=float)
    B = np.broadcast_to(l[None, None, :], shape).copy()
    t.set_fields(b=B, b_next=B)
    t.calc_ext_stim(zz_for(g, g.nz))
    out = t.get_fields()["b"][3, 5]
    c = DT * DB * (g.nz / g.lz) ** 2
    assert out[0] == pytest.approx(c * 2, abs=1e-15)            # mirror: prev_z(0) = 1
    for z in range(1, g.nz - 1):
        assert out[z] == pytest.approx(z - DT * KB * z, rel=1e-14)
    top = g.nz - 1                                               # next_z(NZ-1) = NZ-1
    assert out[top] == pytest.approx(top + c * (-1) - DT * KB * top, rel=1e-14)


def test_b_box_and_untouched_planes(oracle_mod):
    t, g, shape = grid_tissue(oracle_mod)
    rng = np.random.default_rng(1)
    B0 = rng.uniform(0.05, 0.15, shape)
    B1 = rng.uniform(0.05, 0.15, shape)
    t.set_fields(b=B0, b_next=B1)
    izb = 3
    t.calc_ext_stim(zz_for(g, izb))
    f = t.get_fields()
    # after the swap the old "next" buffer is current; planes >= iz_bound keep
    # what that buffer held (double-buffer staleness)
    assert np.array_equal(f["b"][:, :, izb:], B1[:, :, izb:])
    assert np.array_equal(f["b_next"], B0)


def test_flux_weight_is_the_neighbours_mask(oracle_mod):
    t, g, shape = grid_tissue(oracle_mod)
    cm2 = np.ones(shape, np.uint8)
    v = (5, 6, 3)
    cm2[v] = 0
    t.set_maps(cmap1=np.full(shape, -1, np.int32), cmap2=cm2)
    B = np.zeros(shape)
    B[v] = 1.0
    t.set_fields(b=B, b_next=B)
    t.calc_ext_stim(zz_for(g, g.nz))
    out = t.get_fields()["b"]
    c = DT * DB * (g.nz / g.lz) ** 2
    assert out[v] == pytest.approx(1 + c * (-6) - DT * KB, rel=1e-14)
    assert out[4, 6, 3] == 0.0 and out[5, 6, 4] == 0.0          # nothing flows out of v


def test_b_source_window(oracle_mod):
    """fB: +1 on voxels owned by an ALIVE/DEAD cell with 21.5 < agek <= 24."""
    from helpers import sheet_spec
    for agek, on in ((21.4, 0), (21.6, 1), (24.0, 1), (24.1, 0)):
        spec = sheet_spec([dict(state=1, x=8.0, y=8.0, z=2.0, agek=agek)], nmx=4, nmy=4, L=16.0,
                          grid=(16, 16, 8), lz=4.0)
        t = oracle_mod.OracleTissue(spec["geometry"])
        t.load(spec)
        shape = (17, 17, 9)
        c1 = np.full(shape, -1, np.int32)
        c1[8, 8, 4] = spec["nmemb"]
        t.set_maps(cmap1=c1, cmap2=np.ones(shape, np.uint8))
        B = np.zeros(shape)
        t.set_fields(b=B, b_next=B)
        g = spec["geometry"]
        t.calc_ext_stim(zz_for(g, g.nz))
        assert t.get_fields()["b"][8, 8, 4] == pytest.approx(DT * on, abs=1e-18)


@pytest.mark.parametrize("k", [1, 2])
def test_atp_fourier_mode_uses_inv_dx(oracle_mod, k):
    t, g, shape = grid_tissue(oracle_mod)
    j = np.arange(shape[0])
    A = np.broadcast_to(np.cos(2 * np.pi * k * (j % g.nx) / g.nx)[:, None, None], shape).copy()
    t.set_fields(atp=A, atp_next=A)
    t.ca_substep(g.nz)
    out = t.get_fields()["atp"]
    inv_dx2 = (g.nx / g.lx) ** 2
    lam = 1 + DT * (DA * inv_dx2 * (2 * np.cos(2 * np.pi * k / g.nx) - 2) - KAA)
    np.testing.assert_allclose(out[:g.nx, :g.ny, :g.nz], (lam * A)[:g.nx, :g.ny, :g.nz], rtol=0, atol=1e-14)


def test_atp_halo_includes_the_stale_plane(oracle_mod):
    t, g, shape = grid_tissue(oracle_mod)
    rng = np.random.default_rng(2)
    A0 = rng.uniform(0, 1, shape)
    A1 = rng.uniform(0, 1, shape)
    t.set_fields(atp=A0, atp_next=A1)
    izb = 4
    t.ca_substep(izb)
    out = t.get_fields()["atp"]
    # plane iz_bound is not computed but its periodic copies are refreshed
    # from the stale values of the next buffer (ca2p.cpp:334-337)
    assert np.array_equal(out[:g.nx, :g.ny, izb], A1[:g.nx, :g.ny, izb])
    assert np.array_equal(out[:g.nx, g.ny, izb], A1[:g.nx, 0, izb])
    assert np.array_equal(out[g.nx, :g.ny, izb], A1[0, :g.ny, izb])
