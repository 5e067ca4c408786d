"""Triangular membrane bending (cell_interaction.cpp:451-499)."""
import numpy as np
import pytest

from helpers import sheet_spec


def sheet(O, nmx=8, nmy=8, L=8.0, bump=None, scale=1.0):
    spec = sheet_spec([], nmx=nmx, nmy=nmy, L=L * scale, grid=(8, 8, 8), lz=L * scale)
    c = spec["cells"]
    c["x"] *= 1.0  # make_sheet already used the scaled box
    c["z"] = np.full(c["x"].size, 1.0 * scale)
    if bump:
        for m, dz in bump.items():
            c["z"][m] += dz * scale
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    return t, spec


def test_flat_sheet_has_zero_bend(oracle_mod):
    t, spec = sheet(oracle_mod)
    dr = np.array([t.tri_bend(m) for m in range(spec["nmemb"])])
    assert np.all(dr == 0.0)


def test_raised_particle_is_pushed_back(oracle_mod):
    m = 3 * 8 + 3
    t, _ = sheet(oracle_mod, bump={m: 0.05})
    assert t.tri_bend(m)[2] < 0
    t, _ = sheet(oracle_mod, bump={m: -0.05})
    assert t.tri_bend(m)[2] > 0


def test_divisor_is_the_normalised_normal(oracle_mod):
    """tri_bend_1/2 divide by the norm of norm1_pre after normalize() has
    made it a unit vector (vec3::normalize mutates, vec3.h:39-43).  The bend
    term is then homogeneous of degree +1 in the coordinates; with the
    un-normalised norm it would be of degree -1."""
    bump = {3 * 8 + 3: 0.07, 3 * 8 + 4: -0.03, 4 * 8 + 3: 0.02}
    t1, _ = sheet(oracle_mod, bump=bump)
    t2, _ = sheet(oracle_mod, bump=bump, scale=2.0)
    for m in (3 * 8 + 3, 3 * 8 + 4, 2 * 8 + 3):
        a, b = t1.tri_bend(m), t2.tri_bend(m)
        assert b == pytest.approx(2.0 * a, rel=1e-9, abs=1e-15)


def test_bend_is_translation_invariant_across_the_seam(oracle_mod):
    bump = {3 * 8 + 3: 0.07}
    t1, spec = sheet(oracle_mod, bump=bump)
    a = t1.tri_bend(3 * 8 + 3)
    # same configuration shifted by half a box in x (wraps periodically)
    spec2 = sheet_spec([], nmx=8, nmy=8, L=8.0, grid=(8, 8, 8), lz=8.0)
    c = spec2["cells"]
    c["z"] = np.full(c["x"].size, 1.0)
    c["z"][3 * 8 + 3] += 0.07
    c["x"] = np.mod(c["x"] + 4.0, 8.0)
    t2 = oracle_mod.OracleTissue(spec2["geometry"])
    t2.load(spec2)
    assert t2.tri_bend(3 * 8 + 3) == pytest.approx(a, rel=1e-9, abs=1e-15)
