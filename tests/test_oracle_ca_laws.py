"""The per-cell reaction laws of calc_ca2p (ca2p.cpp:36-144) through one
oracle sub-step of an isolated AFM cell, checked against values reduced by
hand at special points of each law.

Special points used (constants from ca2p.cpp):
- ca2p_by_ext_stim(B) = kbc B^2 / (Hb + B^2), kbc = 0.48, Hb = 0.01 (:36-43):
  0 at B = 0; kbc / 2 = 0.24 at B = sqrt(Hb) = 0.1.
- ca2p_into_storage(u) = gamma u / (kg + u), gamma = 2, kg = 0.1 (:45-49):
  0 at u = 0; 1 at u = kg; 1.75 at u = 0.7.
- ER_domain1_active(p) = mu0 + mu1 p / (p + kmu), mu0 = 0.567, mu1 = 0.1,
  kmu = 0.05 (:51-56): 0.567 at p = 0; 0.617 at p = kmu.
- ER_domain2_active(u) = b + bb u / (k1 + u), b = 0.11, bb = 0.89, k1 = 0.7
  (:58-63): 0.11 at u = 0; 0.555 at u = k1; 0.22125 at u = 0.1.
- reaction = k_flux ER1 ER2 h - storage + beta_zero + ext, k_flux = 8.1,
  beta_zero = 0.02 (:72-88).
- calc_th, calc_Kpa, calc_IAG (:91-121) at agek = THRESH_SP (tanh 0): th 0.6,
  Kpa 5, IAG 0.5; at agek = THRESH_SP + 40 (tanh saturates to +-1 in double):
  th 0.2, Kpa 4, IAG 1; non-ALIVE AFM cells: th 1, Kpa 6, IAG 0.
- ex_inert_diff(u, h, th) = (k2^2 / (k2^2 + u^2) - h) / th, k2 = 0.7
  (:123-127): (1 - h) / th at u = 0; (0.5 - h) / th at u = k2; (0.98 - h) / th
  at u = 0.1.
- IP3_default_diff(Kpa, A, p) = Kpa A / (H0 + A) - Kpp p, H0 = 0.5, Kpp = 0.3
  (:129-133, :16): Kpa / 2 - 0.3 p at A = H0.
- grid_avg8 (utils.h:18-23) is the mean of the 8 corners of the cell's voxel:
  on a linear field it equals the field at the voxel centre, so a wrong or
  repeated corner moves it by a multiple of one of the (distinct) gradients.

Each expected value below is the special-point reduction of those laws,
written out per case; a changed constant, a swapped argument or a wrong
corner of the average moves the result far beyond the tolerance.
"""
import numpy as np
import pytest

from helpers import sheet_spec

DT = 0.01
THRESH_SP = 3.0
LAT = (8, 8, 9)         # the cell's voxel: x, y, z in [8, 9), [8, 9), [9, 10)


def linear_field(centre, gx, gy, gz, shape=(17, 17, 17)):
    """centre + gx (j - 8.5) + gy (k - 8.5) + gz (l - 9.5): its value at the
    centre of voxel LAT is `centre`."""
    j, k, l = np.meshgrid(*(np.arange(s, dtype=np.float64) for s in shape), indexing="ij")
    return centre + gx * (j - (LAT[0] + 0.5)) + gy * (k - (LAT[1] + 0.5)) + gz * (l - (LAT[2] + 0.5))


def one_substep(O, state, agek, u, p, h, B, A):
    """One calc_ca2p sub-step of a single AFM cell with no coupled neighbour
    (it floats far above the sheet); returns its (u', p', h')."""
    free = [dict(state=state, x=LAT[0] + 0.3, y=LAT[1] + 0.6, z=LAT[2] + 0.2, agek=agek, ca2p=u, ip3=p)]
    spec = sheet_spec(free, nmx=4, nmy=4, L=16.0, grid=(16, 16, 16))
    t = O.OracleTissue(spec["geometry"])
    t.load(spec)
    m = spec["nmemb"]
    n = t.get_cells()["n"]
    exi = np.full(n, 0.97)
    exi[m] = h
    t.set_cell_values(ex_inert=exi)
    t.set_fields(atp=A, atp_next=A, b=B, b_next=B)
    t.setup_map_lat()
    c = t.get_cells()
    assert tuple(c["lat"][3 * m:3 * m + 3]) == LAT
    assert c["conn_ptr"][m + 1] == c["conn_ptr"][m]            # no neighbour at all
    t.ca_substep(14)
    t.check()
    c = t.get_cells()
    return c["ca2p"][m], c["ip3"][m], c["ex_inert"][m]


def uniform(v):
    return np.full((17, 17, 17), v)


def test_basal_point_u0_p0_B0(oracle_mod):
    # ALIVE at THRESH_SP: th 0.6, Kpa 5.  u = p = B = 0, h = 1, A = H0:
    # reaction = k_flux mu0 b h + beta_zero = 8.1 * 0.567 * 0.11 + 0.02
    u1, p1, h1 = one_substep(oracle_mod, 0, THRESH_SP, 0.0, 0.0, 1.0, uniform(0.0), uniform(0.5))
    assert u1 == pytest.approx(DT * (0.505197 + 0.02), rel=1e-13)
    assert p1 == pytest.approx(DT * 2.5, rel=1e-13)              # Kpa / 2
    assert h1 == 1.0                                            # (1 - h) / th = 0


def test_half_saturation_point(oracle_mod):
    # u = k1 = 0.7, p = kmu = 0.05, B = sqrt(Hb) = 0.1, h = 0.2, A = H0
    # ER1 0.617, ER2 0.555, storage 1.75, ext 0.24:
    # reaction = 8.1 * 0.617 * 0.555 * 0.2 - 1.75 + 0.02 + 0.24 = 0.5547447 - 1.49
    u1, p1, h1 = one_substep(oracle_mod, 0, THRESH_SP, 0.7, 0.05, 0.2, uniform(0.1), uniform(0.5))
    assert u1 == pytest.approx(0.7 + DT * (0.5547447 - 1.49), rel=1e-13)
    assert p1 == pytest.approx(0.05 + DT * (2.5 - 0.015), rel=1e-13)
    assert h1 == pytest.approx(0.2 + DT * (0.5 - 0.2) / 0.6, rel=1e-13)


def test_non_alive_afm_cell(oracle_mod):
    # FIX: th 1, Kpa 6 whatever its agek.  u = kg = 0.1: storage 1, ER2 0.22125,
    # ex_inert_diff (0.98 - h) / 1; p = kmu, B = 0.1, A = H0, h = 1
    for state in (4, 7):                                        # FIX, MUSUME
        u1, p1, h1 = one_substep(oracle_mod, state, THRESH_SP, 0.1, 0.05, 1.0, uniform(0.1), uniform(0.5))
        assert u1 == pytest.approx(0.1 + DT * (8.1 * 0.617 * 0.22125 - 1.0 + 0.02 + 0.24), rel=1e-13)
        assert p1 == pytest.approx(0.05 + DT * (3.0 - 0.015), rel=1e-13)
        assert h1 == pytest.approx(1.0 + DT * (0.98 - 1.0), rel=1e-13)


def test_saturated_age_switches(oracle_mod):
    # ALIVE far past THRESH_SP: th 0.2, Kpa 4; far below: th 1, Kpa 6
    for agek, th, kpa in ((THRESH_SP + 40.0, 0.2, 4.0), (THRESH_SP - 40.0, 1.0, 6.0)):
        u1, p1, h1 = one_substep(oracle_mod, 0, agek, 0.7, 0.05, 0.3, uniform(0.1), uniform(0.5))
        assert u1 == pytest.approx(0.7 + DT * (8.1 * 0.617 * 0.555 * 0.3 - 1.49), rel=1e-13)
        assert p1 == pytest.approx(0.05 + DT * (kpa / 2 - 0.015), rel=1e-13)
        assert h1 == pytest.approx(0.3 + DT * (0.5 - 0.3) / th, rel=1e-13)


def test_grid_avg8_on_linear_fields(oracle_mod):
    # linear B and ATP fields whose voxel-centre values are the special
    # points 0.1 and 0.5: the results equal the uniform case; a wrong corner
    # moves B by >= 0.0007/8 (d ext/dB = 2.4 there) or A by >= 0.001/8
    B = linear_field(0.1, 0.004, 0.0015, 0.0007)
    A = linear_field(0.5, 0.003, -0.002, 0.001)
    u1, p1, h1 = one_substep(oracle_mod, 0, THRESH_SP, 0.7, 0.05, 0.2, B, A)
    assert u1 == pytest.approx(0.7 + DT * (0.5547447 - 1.49), rel=1e-13)
    assert p1 == pytest.approx(0.05 + DT * (2.5 - 0.015), rel=1e-13)
    # and off the special point the average is still the centre value: B = 0.3
    # at the centre gives ext = 0.48 * 0.09 / 0.1 = 0.432
    B3 = linear_field(0.3, -0.005, 0.002, 0.0009)
    u3, _, _ = one_substep(oracle_mod, 0, THRESH_SP, 0.7, 0.05, 0.2, B3, A)
    assert u3 == pytest.approx(0.7 + DT * (0.5547447 - 1.75 + 0.02 + 0.432), rel=1e-13)


def test_grid_avg8_corner_weights(oracle_mod):
    # a single non-zero corner of the cell's voxel contributes 1/8 of its value,
    # and a point just outside the 2x2x2 corner block contributes nothing:
    # the ATP at each corner only enters through Kpa A/(H0 + A)
    base = one_substep(oracle_mod, 0, THRESH_SP, 0.0, 0.0, 1.0, uniform(0.0), uniform(0.0))[1]
    assert base == 0.0
    for corner in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)):
        A = uniform(0.0)
        A[LAT[0] + corner[0], LAT[1] + corner[1], LAT[2] + corner[2]] = 4.0   # mean 0.5
        p1 = one_substep(oracle_mod, 0, THRESH_SP, 0.0, 0.0, 1.0, uniform(0.0), A)[1]
        assert p1 == pytest.approx(DT * 2.5, rel=1e-13), corner
    for outside in ((-1, 0, 0), (2, 0, 0), (0, -1, 0), (0, 2, 0), (0, 0, -1), (0, 0, 2)):
        A = uniform(0.0)
        A[LAT[0] + outside[0], LAT[1] + outside[1], LAT[2] + outside[2]] = 4.0
        p1 = one_substep(oracle_mod, 0, THRESH_SP, 0.0, 0.0, 1.0, uniform(0.0), A)[1]
        assert p1 == 0.0, outside


def test_iag_scales_the_exchange(oracle_mod):
    # two ALIVE cells in contact past THRESH_SP + 40: IAG = 1, so the Ca2+ and
    # IP3 exchanges are du * gj * (u_o - u) and dp * gj * (p_o - p) exactly
    # (ca2p.cpp:199-216; gj = gj_init = 0.99 on both directed edges)
    def pair(far):
        free = [dict(state=0, x=8.0, y=8.0, z=9.0, agek=43.0, ca2p=0.10, ip3=0.20),
                dict(state=0, x=2.0 if far else 10.5, y=2.0 if far else 8.0, z=9.0, agek=43.0,
                     ca2p=0.30, ip3=0.05)]
        spec = sheet_spec(free, nmx=4, nmy=4, L=16.0, grid=(16, 16, 16))
        t = oracle_mod.OracleTissue(spec["geometry"])
        t.load(spec)
        t.set_fields(atp=uniform(0.1), atp_next=uniform(0.1), b=uniform(0.1), b_next=uniform(0.1))
        t.setup_map_lat()
        t.ca_substep(12)
        return t.get_cells(), spec["nmemb"]
    (cp, m), (ci, _) = pair(False), pair(True)
    ex_u = (cp["ca2p"][m] - ci["ca2p"][m]) / DT
    ex_p = (cp["ip3"][m] - ci["ip3"][m]) / DT
    assert ex_u == pytest.approx(0.01 * 1.0 * 0.99 * 0.20, rel=1e-9)
    assert ex_p == pytest.approx(0.1 * 1.0 * 0.99 * -0.15, rel=1e-9)


def test_non_alive_cells_take_no_exchange(oracle_mod):
    # calc_IAG is iage_kitei = 0 for FIX and MUSUME (ca2p.cpp:113-121): a
    # MUSUME coupled to an ALIVE cell keeps its isolated update, while the
    # ALIVE side (agek = THRESH_SP + 40, IAG = 1) takes du * gj * (u_o - u)
    def pair(far):
        free = [dict(state=7, x=8.0, y=8.0, z=9.0, ca2p=0.10, ip3=0.20),
                dict(state=0, x=2.0 if far else 10.5, y=2.0 if far else 8.0, z=9.0, agek=43.0,
                     ca2p=0.30, ip3=0.05)]
        spec = sheet_spec(free, nmx=4, nmy=4, L=16.0, grid=(16, 16, 16))
        t = oracle_mod.OracleTissue(spec["geometry"])
        t.load(spec)
        t.set_fields(atp=uniform(0.1), atp_next=uniform(0.1), b=uniform(0.1), b_next=uniform(0.1))
        t.setup_map_lat()
        c = t.get_cells()
        m = spec["nmemb"]
        coupled = m + 1 in c["conn"][c["conn_ptr"][m]:c["conn_ptr"][m + 1]]
        assert coupled != far
        t.ca_substep(12)
        return t.get_cells(), m
    (cp, m), (ci, _) = pair(False), pair(True)
    assert cp["ca2p"][m] == ci["ca2p"][m] and cp["ip3"][m] == ci["ip3"][m]
    assert (cp["ca2p"][m + 1] - ci["ca2p"][m + 1]) / DT == pytest.approx(0.01 * 0.99 * -0.20, rel=1e-9)
