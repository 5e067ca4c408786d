"""The C++ facade (include/epi_amd.hpp) and the driver loop written against it
(epi_amd/host/proc.cpp, the reference's proc.cpp:131-173 call for call).

CPU: the header compiles on its own, the deck format round-trips, the driver
parses its arguments.  GPU: the driver, fed a deck through cman_init,
reaches the same state as the Python driver of the same library through
epi_step (bit-exact), and that state agrees with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from epi_amd import configs, deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "epi_amd", "host", "epi_proc")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "epi_amd.hpp"\n'
                   "int main() { epi_amd::CellManager* m = nullptr; (void)m; return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        f"-I{os.path.join(ROOT, 'include')}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_deck_round_trip(tmp_path):
    spec = configs.config_tiny(7, 160)
    p = tmp_path / "deck.txt"
    deck.write_deck(spec, p)
    back = deck.spec_from_columns(deck.parse_deck(p), spec["geometry"])
    assert back["nmemb"] == spec["nmemb"] and back["nder"] == spec["nder"]
    for k in ("x", "y", "z", "radius", "agek", "ageb", "ex_fat", "in_fat", "spr_nat_len"):
        assert np.array_equal(back["cells"][k], spec["cells"][k]), k
    for k in ("state", "pair", "rest_div_times", "fix_origin"):
        assert np.array_equal(back["cells"][k], spec["cells"][k]), k
    # the loader's flag: stem_orig_id < MALIG_NUM = 0 (cellmanager.cpp:292)
    assert np.array_equal(back["cells"]["is_malignant"], (spec["cells"]["fix_origin"] < 0).astype(np.uint8))


def test_driver_usage():
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def _geometry_args(g):
    return ["--geometry", str(g.nmx), str(g.nmy), str(g.nx), str(g.ny), str(g.nz), repr(g.lx),
            repr(g.ly), repr(g.lz), str(g.max_cells)]


@pytest.mark.gpu
def test_driver_matches_python_driver_and_oracle(tmp_path, oracle_mod):
    from epi_amd import native
    assert os.path.exists(DRIVER), "epi_amd/host/epi_proc is not built"
    spec0 = configs.config_tiny(7, 160)
    g = spec0["geometry"]
    p = tmp_path / "deck.txt"
    deck.write_deck(spec0, p)
    steps, ca_every, seed = 40, 20, 11
    dump = tmp_path / "dump.bin"
    snaps = tmp_path / "snaps"
    snaps.mkdir()
    print(f"facade driver: tiny deck, {steps} steps, Ca every {ca_every}, seed {seed}", flush=True)
    r = subprocess.run([DRIVER, str(p), str(steps), "--seed", str(seed), "--ca-every", str(ca_every),
                        "--dump", str(dump), "--snapshots", str(snaps), "--cut", "10", *_geometry_args(g)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the numbered series the reference's viewer reads: steps 0, 10, 20, 30,
    # each line parsed by its sscanf layout (epi_vis_lightweight/main.cpp:521)
    assert sorted(os.listdir(snaps)) == ["0", "1", "2", "3"]
    rows = [ln.split() for ln in (snaps / "0").read_text().splitlines()]
    assert rows and all(len(r_) == 17 for r_ in rows)
    state = np.array([int(r_[1]) for r_ in rows])
    xyz0 = np.array([[float(r_[6]), float(r_[7]), float(r_[8])] for r_ in rows])
    assert np.array_equal(state, spec0["cells"]["state"])
    # CellManager::output prints doubles with 15 decimals (fixed << setprecision(15))
    assert np.allclose(xyz0, np.stack([spec0["cells"][k] for k in ("x", "y", "z")], 1), rtol=0, atol=1e-14)
    raw = np.fromfile(dump, dtype=np.float64)
    n = int(raw[0])
    x, y, z, ca2p, agek = raw[1:].reshape(5, n)

    spec = deck.spec_from_columns(deck.parse_deck(p), g)
    draws = native.mt19937_64_uniform(seed, 200000)
    t = native.Tissue(g)
    t.load(spec, value_init=True)
    t.set_rng(draws)
    drv = native.Driver(0, 0, 0, ca_every, 0)
    t.step(drv, 30)
    # snapshot 3 (before step 30), taken asynchronously by the driver: the
    # text output() writes from a synchronous download at that step
    c30 = t.get_cells()
    st, ptr, conn = c30["state"], c30["conn_ptr"], c30["conn"]
    touch = [int(st[i] == configs.ALIVE and np.any(st[conn[ptr[i]:ptr[i + 1]]] == configs.DEAD))
             for i in range(st.size)]
    want = "".join("%d %d %.15f %.15f %.15f %.15f %.15f %.15f %.15f %.15f %d %.15f %.15f %d %.15f %d %d\n" % (
        i, st[i], c30["radius"][i], c30["ageb"][i], c30["agek"][i], c30["ca2p"][i], c30["x"][i], c30["y"][i],
        c30["z"][i], c30["ca2p_avg"][i], c30["rest_div_times"][i], c30["ex_fat"][i], c30["in_fat"][i], touch[i],
        c30["spr_nat_len"][i], c30["pair"][i], c30["fix_origin"][i]) for i in range(st.size))
    assert (snaps / "3").read_text() == want
    t.step(drv, steps - 30)
    assert t.error() == 0
    c = t.get_cells()
    assert c["x"].size == n
    for k, v in (("x", x), ("y", y), ("z", z), ("ca2p", ca2p), ("agek", agek)):
        assert np.array_equal(c[k], v), k

    o = oracle_mod.OracleTissue(g)
    o.load(spec, value_init=True)
    o.set_rng(draws)
    drv = oracle_mod.Driver(0, 0, 0, ca_every, 0)
    for _ in range(steps):
        o.step(drv)
    assert o.error() == 0
    oc = o.get_cells()
    assert oc["x"].size == n
    for k, v in (("x", x), ("y", y), ("z", z)):
        assert np.array_equal(oc[k], v), k
    # Ca values carry ulp differences of tanh between glibc and the GPU
    rel = np.max(np.abs(oc["ca2p"] - ca2p) / np.maximum(np.abs(oc["ca2p"]), 1e-300))
    assert rel <= 1e-10, rel


REC = np.dtype([("index", "<i4"), ("state", "<i4"), ("radius", "<f8"), ("ageb", "<f8"), ("agek", "<f8"),
                ("ca2p", "<f8"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("ca2p_avg", "<f8"),
                ("rest_div_times", "<i4"), ("ex_fat", "<f8"), ("in_fat", "<f8"), ("is_touch", "i1"),
                ("spr_nat_len", "<f8"), ("pair", "<i4"), ("fix_origin", "<i4")])


def test_binary_record_layout():
    """CellManager::output's binary record (cellmanager.cpp:496-513): the
    fields back to back, no padding, 109 bytes"""
    assert REC.itemsize == 4 + 4 + 8 * 8 + 4 + 8 + 8 + 1 + 8 + 4 + 4 == 109


def _read_w(path):
    out, lines = {}, open(path).read().splitlines()
    i = 0
    while i < len(lines):
        c, k = map(int, lines[i].split())
        out[c] = [(int(a), b) for a, b in (ln.split() for ln in lines[i + 1:i + 1 + k])]
        i += 1 + k
    return out


@pytest.mark.gpu
def test_last_data_round_trip(tmp_path):
    """output_field_data's five files (proc.cpp:74-84) and the binary form of
    CellManager::output, written by the driver after a run, read back by
    cman_init with use_last and Field::read_binary (proc.cpp:98-110), and
    written again: the field binaries and the uvp / w text come back
    unchanged, the cells within the text's 15 decimals, and the binary
    record equals the run's final state bit for bit."""
    assert os.path.exists(DRIVER), "epi_amd/host/epi_proc is not built"
    spec0 = configs.config_tiny(7, 160)
    g = spec0["geometry"]
    deck_p = tmp_path / "deck.txt"
    deck.write_deck(spec0, deck_p)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    dump, binp = tmp_path / "dump.bin", tmp_path / "cells.bin"
    print("last data round trip: 30 steps, Ca every 15, then a restart", flush=True)
    r = subprocess.run([DRIVER, str(deck_p), "30", "--seed", "5", "--ca-every", "15", "--save-last", str(a),
                        "--output-binary", str(binp), "--dump", str(dump), *_geometry_args(g)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    names = ["last_data_cell", "last_data_uvp", "last_data_w_alt", "last_data_a_alt", "last_data_B_alt"]
    assert sorted(os.listdir(a)) == sorted(names)
    raw = np.fromfile(dump, dtype=np.float64)
    n = int(raw[0])
    x, y, z, ca2p, agek = raw[1:].reshape(5, n)
    rec = np.fromfile(binp, dtype=REC)
    assert rec.size == n and np.array_equal(rec["index"], np.arange(n))
    for k, v in (("x", x), ("y", y), ("z", z), ("ca2p", ca2p), ("agek", agek)):
        assert np.array_equal(rec[k], v), k
    # is_touch: ALIVE with a DEAD neighbour (check_localization); the text
    # columns agree with the binary record
    txt = np.loadtxt(a / "last_data_cell", ndmin=2)
    assert txt.shape == (n, 17)
    for col, k in ((1, "state"), (10, "rest_div_times"), (13, "is_touch"), (15, "pair"), (16, "fix_origin")):
        assert np.array_equal(txt[:, col].astype(np.int64), rec[k].astype(np.int64)), k
    for col, k in ((2, "radius"), (3, "ageb"), (4, "agek"), (5, "ca2p"), (6, "x"), (7, "y"), (8, "z"),
                   (9, "ca2p_avg"), (11, "ex_fat"), (12, "in_fat"), (14, "spr_nat_len")):
        assert np.allclose(txt[:, col], rec[k], rtol=0, atol=1e-14), k
    # restart from the saved set and save again at once
    r = subprocess.run([DRIVER, str(a / "last_data_cell"), "0", "--use-last", str(a), "--save-last", str(b),
                        *_geometry_args(g)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for k in ("last_data_a_alt", "last_data_B_alt"):
        fa, fb = (a / k).read_bytes(), (b / k).read_bytes()
        assert len(fa) == 8 * (g.nx + 1) * (g.ny + 1) * (g.nz + 1) and fa == fb, k
    # 15 significant digits read and written again give the same text
    assert (a / "last_data_uvp").read_text() == (b / "last_data_uvp").read_text()
    wa, wb = _read_w(a / "last_data_w_alt"), _read_w(b / "last_data_w_alt")
    assert len(wa) == n and all(sorted(wa[c]) == sorted(wb[c]) for c in wa)
    ta, tb = np.loadtxt(a / "last_data_cell", ndmin=2), np.loadtxt(b / "last_data_cell", ndmin=2)
    assert np.array_equal(ta[:, [0, 1, 10, 15, 16]], tb[:, [0, 1, 10, 15, 16]])
    assert np.allclose(ta, tb, rtol=0, atol=1e-14)


@pytest.mark.gpu
def test_driver_runs_under_a_parameter_file(tmp_path):
    """`epi_proc --params FILE` (a CalcParams key=value file, every key
    present as Params::load requires) runs the loop under those values, the
    same bits as the Python driver given the same record; `--params-out`
    writes the record back"""
    from epi_amd import native
    assert os.path.exists(DRIVER), "epi_amd/host/epi_proc is not built"
    spec0 = configs.config_tiny(7, 160)
    g = spec0["geometry"]
    p = tmp_path / "deck.txt"
    deck.write_deck(spec0, p)
    prm = native.params_default().replace(LX=g.lx, LY=g.ly, LZ=g.lz, NX=g.nx, NY=g.ny, NZ=g.nz, DT_Cell=0.0105,
                                          K_TOTAL=3.2, kpa=4.3, SW_THRESH=10, Ca_avg_time=3.09, eps_kb=0.13)
    pf, pf2 = tmp_path / "prm", tmp_path / "prm_out"
    native.params_save(str(pf), prm)
    steps, ca_every, seed = 40, 20, 11
    dump = tmp_path / "dump.bin"
    print(f"facade driver under a parameter file: tiny deck, {steps} steps", flush=True)
    r = subprocess.run([DRIVER, str(p), str(steps), "--seed", str(seed), "--ca-every", str(ca_every), "--dump", str(dump),
                        "--params", str(pf), "--params-out", str(pf2), *_geometry_args(g)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert native.params_load(str(pf2)).to_dict() == prm.to_dict()
    raw = np.fromfile(dump, dtype=np.float64)
    n = int(raw[0])
    x, y, z, ca2p, agek = raw[1:].reshape(5, n)
    spec = deck.spec_from_columns(deck.parse_deck(p), g)
    t = native.Tissue(g)
    t.set_params(prm)
    t.load(spec, value_init=True)
    t.set_rng(native.mt19937_64_uniform(seed, 200000))
    t.step(native.Driver(0, 0, 0, ca_every, 0), steps)
    assert t.error() == 0
    c = t.get_cells()
    for k, v in (("x", x), ("y", y), ("z", z), ("ca2p", ca2p), ("agek", agek)):
        assert np.array_equal(c[k], v), k
