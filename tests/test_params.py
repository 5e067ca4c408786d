"""Runtime parameters: the key=value reader and writer (epi_intg/parser.cpp:
30-58, epi_intg/Params.cpp:16-110), epi5_MT's defaults, and the oracle's
restatement of them.  Host code only (no GPU call).

The reference's own parameter file (epi_intg/ppsss, kept as the fixture
tests/golden/ppsss.json) pins the defaults: it lists every key with
epi5_MT's value printed to six decimals."""
import json
import math
import os

import numpy as np
import pytest

from epi_amd import native as N

HERE = os.path.dirname(os.path.abspath(__file__))
PPSSS = json.load(open(os.path.join(HERE, "golden", "ppsss.json")))["pairs"]
INT_KEYS = {k for k, t in N.Params._fields_ if t is N.C.c_int32}


def write(path, lines):
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return str(path)


def ppsss_file(tmp_path, name="ppsss"):
    return write(tmp_path / name, [f"{k}={v}" for k, v in PPSSS])


def test_header_layout_matches_both_bindings(oracle_mod):
    """the product binding, the oracle binding and the C header agree on the
    record's fields, types and size"""
    hdr = oracle_mod.header_param_fields()
    assert [(k, t) for k, t in N.Params._fields_][:len(hdr)] == hdr
    assert len(N.Params._fields_) == len(hdr) == 110
    assert N.C.sizeof(N.Params) == N.C.sizeof(oracle_mod.Params)


def test_fixture_lists_every_key_once():
    keys = [k for k, _ in PPSSS]
    assert len(keys) == len(set(keys))
    assert set(keys) == set(N.PARAM_KEYS)


def test_defaults_agree_with_the_reference_parameter_file(tmp_path):
    """every key of epi_intg/ppsss, read by the library's reader, equals the
    default to the file's six printed decimals (kbc = 0.4*1.2 and S0 =
    0.1*0.2 are products in epi5_MT, so the printed value differs in the last
    bit), and outputdir is the file's"""
    d = N.params_default().to_dict()
    p = N.params_load(ppsss_file(tmp_path), require_all=True).to_dict()
    for k, v in PPSSS:
        if k == "outputdir":
            assert p[k] == "output_test"
            continue
        if k in INT_KEYS:
            assert p[k] == d[k] == int(float(v)), k
        else:
            assert p[k] == float(v), k
            assert abs(d[k] - float(v)) <= 5e-7 * max(1.0, abs(d[k])), (k, d[k], v)
            assert d[k] == float(v) or k in ("kbc", "S0"), k


def test_defaults_match_the_oracle_restatement(oracle_mod):
    d = N.params_default().to_dict()
    o = oracle_mod.params_default().to_dict()
    assert d == o
    # spot values from the reference's sources, independent of both
    assert d["DT_Cell"] == 0.01 and d["DT_Ca"] == 0.01                 # define.h:86-89
    assert d["kbc"] == 0.4 * 1.2 and d["S0"] == 0.1 * 0.2              # ca2p.cpp:10, cell_state_renew.cpp
    assert d["THRESH_DEAD"] == 22.0 and d["ADHE_CONST"] == 31.3
    assert d["N3"] == 200 and d["MAX_CONNECT_CELL_NUM"] == 400


def test_save_then_load_is_exact(tmp_path):
    """epi_params_save writes 17 significant digits, so any record reads back
    bit for bit (the reference writes six decimals)"""
    rng = np.random.default_rng(3)
    p = N.params_default()
    for k, t in N.Params._fields_:
        if t is N.C.c_double:
            setattr(p, k, float(getattr(p, k)) * (1 + 0.37 * rng.uniform(-1, 1)) + 1e-3 * rng.uniform())
        elif t is N.C.c_int32 and k not in ("USE_TRI_MEMB", "NEW_BEND_POT", "STOCHASTIC", "force_cornification"):
            setattr(p, k, int(getattr(p, k)) + int(rng.integers(0, 50)))
    p.outputdir = b"out dir/x"
    path = str(tmp_path / "saved")
    N.params_save(path, p)
    q = N.params_load(path, require_all=True, base=N.Params())
    assert q.to_dict() == p.to_dict()
    assert len(open(path).read().splitlines()) == 110


def test_written_order_is_calcparams_order(tmp_path):
    """generate_paramfile writes piset's order (CalcParams.cpp:51-162); the
    fixture file lists outputdir before the last two bools, the writer after"""
    path = str(tmp_path / "saved")
    N.params_save(path, N.params_default())
    keys = [ln.split("=", 1)[0] for ln in open(path).read().splitlines()]
    fixture = [k for k, _ in PPSSS]
    assert keys[-3:] == ["STOCHASTIC", "force_cornification", "outputdir"]
    assert [k for k in keys if k != "outputdir"] == [k for k in fixture if k != "outputdir"]


def test_parser_rules(tmp_path):
    """parse_paramtext (parser.cpp:30-58): '#' and '//' at column 0 comment a
    line out, a line without '=' is skipped, spaces around key and value are
    dropped, the first line of a key wins; aliases of CalcParams::init"""
    path = write(tmp_path / "p", [
        "# DT_Cell=5",
        "// DT_Ca=5",
        "no separator here",
        "   DT_Cell   =   0.02   ",
        "DT_Cell=0.03",
        "LOOP=123",
        "MALIGNANT=4",
        "Ca2P_avg_time=5.5",
        "TURNOVER=77",
        "STOCHASTIC=2",
        "NEW_BEND_POT=1",
        "outputdir='a  b'",
        "unknown_key=1",
        "DISA_conn_num_thresh=11.000000",
        "SW_THRESH=7abc",
    ])
    p = N.params_load(path, require_all=False)
    assert p.DT_Cell == 0.02 and p.DT_Ca == 0.01
    assert p.NUM_ITR == 123 and p.MALIG_NUM == 4 and p.Ca_avg_time == 5.5 and p.T_TURNOVER == 77
    assert p.STOCHASTIC == 0 and p.NEW_BEND_POT == 1          # a bool is "== 1" (Params.cpp:39-41)
    assert p.outputdir == b"a b"                               # remove_LR_space folds inner runs
    assert p.DISA_conn_num_thresh == 11.0
    assert p.SW_THRESH == 7                                    # stoul reads the leading number


@pytest.mark.parametrize("line,msg", [
    ("DT_Cell=abc", "DT_Cell"),
    ("NX=-5", "NX"),
    ("outputdir=noquotes", "outputdir"),
])
def test_bad_values_are_errors(tmp_path, line, msg):
    path = write(tmp_path / "p", [line])
    with pytest.raises(N.EpiError) as e:
        N.params_load(path, require_all=False)
    assert e.value.code == 7 and msg in str(e.value)


def test_require_all_names_the_missing_key(tmp_path):
    lines = [f"{k}={v}" for k, v in PPSSS if k != "kpa"]
    path = write(tmp_path / "p", lines)
    with pytest.raises(N.EpiError) as e:
        N.params_load(path, require_all=True)
    assert "kpa must be specified." in str(e.value)
    p = N.params_load(path, require_all=False)
    assert p.kpa == 4.0


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(N.EpiError) as e:
        N.params_load(str(tmp_path / "absent"))
    assert "Parse failed." in str(e.value)


def test_oracle_tissue_parameters(oracle_mod):
    """a new oracle tissue holds the defaults with its geometry's grid and
    box, and reports back what it was given"""
    from epi_amd import configs
    g = configs.config_tiny(3, n_free=50, nz=24)["geometry"]
    o = oracle_mod.OracleTissue(g)
    d = o.get_params().to_dict()
    assert d["NX"] == g.nx and d["LZ"] == g.lz
    assert {k: v for k, v in d.items() if k not in ("LX", "LY", "LZ", "NX", "NY", "NZ")} == \
        {k: v for k, v in N.params_default().to_dict().items() if k not in ("LX", "LY", "LZ", "NX", "NY", "NZ")}
    o.set_params({"DT_Ca": 0.02, "Ca_avg_time": 5.0})
    q = o.get_params()
    assert q.DT_Ca == 0.02 and q.Ca_avg_time == 5.0 and math.isclose(q.Ca_avg_time / q.DT_Ca, 250)
