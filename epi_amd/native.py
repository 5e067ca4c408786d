"""ctypes binding of the HIP library (include/epi_amd.h).

The package has no CPU fallback: if libepi_amd.so is missing the import of
this module fails with the build command to run.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# EPI_AMD_LIB names another build of the same library (to compare builds)
LIB_PATH = os.environ.get("EPI_AMD_LIB") or os.path.join(HERE, "libepi_amd.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)

ERRORS = {0: "OK", 1: "CAPACITY", 2: "STRONG_INTERACTION", 3: "NO_DERMIS",
          4: "OUT_OF_GRID", 5: "REMOVE_ALIAS", 6: "RNG_EXHAUSTED", 7: "INVALID",
          8: "HIP", 9: "PAIR_REMOVED"}


class EpiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code


class Geometry(C.Structure):
    _fields_ = [("nmx", C.c_int32), ("nmy", C.c_int32), ("nx", C.c_int32),
                ("ny", C.c_int32), ("nz", C.c_int32), ("lx", C.c_double),
                ("ly", C.c_double), ("lz", C.c_double), ("max_cells", C.c_int32),
                ("row_stride", C.c_int32), ("bucket_cap", C.c_int32)]


class CellInput(C.Structure):
    _fields_ = [("n", C.c_int32), ("nmemb", C.c_int32), ("nder", C.c_int32),
                ("state", _u8p), ("fix_origin", _ip),
                ("x", _dp), ("y", _dp), ("z", _dp), ("radius", _dp), ("ca2p", _dp),
                ("ca2p_avg", _dp), ("ip3", _dp), ("ex_inert", _dp),
                ("agek", _dp), ("ageb", _dp), ("ex_fat", _dp), ("in_fat", _dp),
                ("spr_nat_len", _dp), ("rest_div_times", _ip), ("is_malignant", _u8p),
                ("pair", _ip)]


class Cells(C.Structure):
    _fields_ = [("n", C.c_int32), ("nmemb", C.c_int32), ("nder", C.c_int32),
                ("sw", C.c_int32),
                ("state", _u8p), ("fix_origin", _ip),
                ("x", _dp), ("y", _dp), ("z", _dp), ("xn", _dp), ("yn", _dp), ("zn", _dp),
                ("radius", _dp),
                ("ca2p", _dp), ("ca2p_next", _dp), ("ip3", _dp), ("ip3_next", _dp),
                ("ca2p_avg", _dp), ("ex_inert", _dp), ("agek", _dp), ("ageb", _dp),
                ("ex_fat", _dp), ("in_fat", _dp), ("spr_nat_len", _dp), ("diff_u", _dp),
                ("rest_div_times", _ip), ("pair", _ip), ("dermis", _ip), ("lat", _ip),
                ("is_malignant", _u8p), ("uid", _i64p),
                ("conn_ptr", _ip), ("conn", _ip), ("gj", _dp), ("conn_cap", C.c_int32),
                ("memb_adj", _ip), ("cst", _i64p)]


class Driver(C.Structure):
    _fields_ = [("step", C.c_int64), ("force_cornif", C.c_int32), ("num_sc", C.c_int32),
                ("ca_every", C.c_int32), ("clean_every", C.c_int32)]


SNAPSHOT_F64 = ("radius", "ageb", "agek", "ca2p", "x", "y", "z", "ca2p_avg", "ex_fat", "in_fat", "spr_nat_len")


class Snapshot(C.Structure):
    """epi_snapshot: CellManager::output's per-cell record, downloaded
    asynchronously (pointers into the context's pinned memory)"""
    _fields_ = [("tissue", C.c_int32), ("n", C.c_int32), ("state", C.POINTER(C.c_uint8)),
                ("touch", C.POINTER(C.c_uint8)), ("rest_div_times", C.POINTER(C.c_int32)),
                ("pair", C.POINTER(C.c_int32)), ("fix_origin", C.POINTER(C.c_int32))] + \
               [(k, C.POINTER(C.c_double)) for k in SNAPSHOT_F64]


class Status(C.Structure):
    _fields_ = [("n", C.c_int32), ("err", C.c_int32), ("sw", C.c_int32), ("ndiv", C.c_int32),
                ("nrm", C.c_int32), ("maxlen", C.c_int32), ("rng_pos", C.c_int64),
                ("next_uid", C.c_int64), ("zzmax", C.c_double), ("ca_solves", C.c_int64),
                ("steps", C.c_int64), ("cell_steps", C.c_int64), ("grid_points_b", C.c_int64),
                ("grid_points_ca", C.c_int64)]


CELL_DTYPES = {
    "state": np.uint8, "fix_origin": np.int32, "x": np.float64, "y": np.float64,
    "z": np.float64, "xn": np.float64, "yn": np.float64, "zn": np.float64,
    "radius": np.float64, "ca2p": np.float64, "ca2p_next": np.float64,
    "ip3": np.float64, "ip3_next": np.float64, "ca2p_avg": np.float64,
    "ex_inert": np.float64, "agek": np.float64, "ageb": np.float64,
    "ex_fat": np.float64, "in_fat": np.float64, "spr_nat_len": np.float64,
    "diff_u": np.float64, "rest_div_times": np.int32, "pair": np.int32,
    "dermis": np.int32, "is_malignant": np.uint8, "uid": np.int64,
}

SYMBOLS = [
    "epi_create", "epi_destroy", "epi_last_error", "epi_num_tissues", "epi_load_cells",
    "epi_value_init", "epi_set_cell_values", "epi_set_fields", "epi_set_rng",
    "epi_rng_mt19937_64", "epi_cell_interaction", "epi_pos_copy", "epi_cell_state_renew",
    "epi_cell_pos_periodic_fix", "epi_connect_cell", "epi_calc_zzmax", "epi_setup_map_lat",
    "epi_calc_ext_stim", "epi_initialize_sc", "epi_calc_ca2p", "epi_should_calc_ca",
    "epi_ca_calc_condition_reset", "epi_ca_substep", "epi_step", "epi_synchronize",
    "epi_get_status", "epi_get_cells", "epi_get_fields", "epi_set_positions",
    "epi_set_state", "epi_set_lists", "epi_set_maps", "epi_profile_enable",
    "epi_profile_read", "epi_profile_reset", "epi_profile_kernel_count",
    "epi_set_tuning", "epi_get_tuning", "epi_calc_ca2p_n", "epi_profile_sample", "epi_profile_launches",
    "epi_rd_begin", "epi_rd_forward", "epi_rd_get_ca2p", "epi_rd_adjoint", "epi_rd_end",
    "epi_mech_begin", "epi_mech_forward", "epi_mech_get_positions", "epi_mech_adjoint", "epi_mech_end",
    "epi_set_timestep", "epi_take_cell_stats", "epi_params_default", "epi_params_load", "epi_params_save",
    "epi_set_params", "epi_get_params", "epi_sheet_size", "epi_init_gen",
    "epi_snapshot_begin", "epi_snapshot_wait",
]

SNAPSHOT_SLOTS = 2   # EPI_SNAPSHOT_SLOTS

MECH_DEFAULT = {"eps": 0.01, "k_adh": 3.0, "k_spring": 25.0, "k_bend": 0.5}

# launch-shape keys of epi_set_tuning (include/epi_amd.h)
TUNING_KEYS = {"b_zb": 0, "b_rows": 1, "atp_zb": 2, "atp_rows": 3, "ca_group": 4, "xcd": 5, "rd_group": 6, "ca_fuse": 7,
               "map_lanes": 8, "ca_group_small": 9, "step_graph": 10, "b_xcd": 11,
               "sweep_wave": 12, "side_fork": 13}

class Params(C.Structure):
    """epi_params (include/epi_amd.h): the keys of epi_intg's CalcParams
    (epi_intg/calc/CalcParams.cpp:51-162) in the header's order."""
    _fields_ = [
        ("MAX_CONNECT_CELL_NUM", C.c_int32), ("DT_Cell", C.c_double), ("DT_Ca", C.c_double),
        ("MALIG_NUM", C.c_int32), ("LX", C.c_double), ("LY", C.c_double), ("LZ", C.c_double), ("NX", C.c_int32),
        ("NY", C.c_int32), ("NZ", C.c_int32), ("R_max", C.c_double), ("R_der", C.c_double), ("R_memb", C.c_double),
        ("COMPRESS_FACTOR", C.c_double), ("MEMB_ADHE_RANGE", C.c_int32), ("THRESH_SP", C.c_double),
        ("LJ_THRESH", C.c_double), ("THRESH_DEAD", C.c_double), ("NUM_ITR", C.c_int32), ("Ca_avg_time", C.c_double),
        ("ca2p_init", C.c_double), ("ex_inert_init", C.c_double), ("IP3_init", C.c_double), ("gj_init", C.c_double),
        ("ATP_init", C.c_double), ("ext_stim_init", C.c_double), ("FAC_MAP", C.c_double), ("SW_THRESH", C.c_int32),
        ("T_TURNOVER", C.c_double), ("NUM_SC_INIT", C.c_int32), ("USE_TRI_MEMB", C.c_int32),
        ("agki_max", C.c_double), ("fac", C.c_double), ("div_max", C.c_int32), ("Kpp", C.c_double),
        ("dp", C.c_double), ("kbc", C.c_double), ("Hb", C.c_double), ("Cout", C.c_double), ("kg", C.c_double),
        ("gamma", C.c_double), ("mu0", C.c_double), ("mu1", C.c_double), ("kmu", C.c_double), ("para_b", C.c_double),
        ("para_bb", C.c_double), ("para_k1", C.c_double), ("k_flux", C.c_double), ("beta_zero", C.c_double),
        ("CA_OUT", C.c_double), ("thgra", C.c_double), ("delta_th", C.c_double), ("thpri", C.c_double),
        ("kpa", C.c_double), ("delta_K", C.c_double), ("Kpri", C.c_double), ("delta_I", C.c_double),
        ("iage_kitei", C.c_double), ("para_k2", C.c_double), ("H0", C.c_double), ("wd", C.c_double),
        ("ca2p_du", C.c_double), ("STIM11", C.c_double), ("Kaa", C.c_double), ("Da", C.c_double),
        ("AIR_STIM", C.c_double), ("AREA_GRID_ORIGINAL", C.c_double), ("N2", C.c_int32), ("N3", C.c_int32),
        ("eps_m", C.c_double), ("DER_DER_CONST", C.c_double), ("K_TOTAL", C.c_double),
        ("K_DESMOSOME_RATIO", C.c_double), ("Kspring", C.c_double), ("Kspring_d", C.c_double), ("KBEND", C.c_double),
        ("delta_R_coef", C.c_double), ("para_ljp2", C.c_double), ("Kspring_division", C.c_double),
        ("stoch_div_time_ratio", C.c_double), ("S2", C.c_double), ("accel_div", C.c_double), ("eps_kb", C.c_double),
        ("alpha_b", C.c_double), ("S0", C.c_double), ("eps_ks", C.c_double), ("accel_diff", C.c_double),
        ("alpha_k", C.c_double), ("eps_kk", C.c_double), ("S1", C.c_double), ("ubar", C.c_double),
        ("delta_lipid", C.c_double), ("delta_sig_r1", C.c_double), ("lipid_rel", C.c_double), ("lipid", C.c_double),
        ("delta_L_coef", C.c_double), ("ADHE_CONST", C.c_double), ("DISA_conn_num_thresh", C.c_double),
        ("eps_L", C.c_double), ("unpair_dist_coef", C.c_double), ("kb", C.c_double), ("DUR_ALIVE", C.c_double),
        ("DUR_DEAD", C.c_double), ("DB", C.c_double), ("SYSTEM", C.c_int32), ("CUT", C.c_int32),
        ("NEW_BEND_POT", C.c_int32), ("STOCHASTIC", C.c_int32), ("force_cornification", C.c_int32),
        ("outputdir", C.c_char * 256),
    ]

    def to_dict(self):
        return {k: (getattr(self, k).decode() if k == "outputdir" else getattr(self, k)) for k, _ in self._fields_}

    def replace(self, **kv):
        """a copy with the given keys changed"""
        q = Params()
        C.pointer(q)[0] = self
        for k, v in kv.items():
            if k not in PARAM_KEYS:
                raise KeyError(k)
            setattr(q, k, v.encode() if k == "outputdir" else v)
        return q


PARAM_KEYS = [k for k, _ in Params._fields_]


def params_default():
    """epi5_MT's values (epi_params_default)"""
    p = Params()
    load_library().epi_params_default(C.byref(p))
    return p


def params_load(path, require_all=True, base=None):
    """a ppsss-style key=value file (Params::load over parse_paramtext,
    epi_intg/Params.cpp:91-96, parser.cpp:30-58) read over `base` (default:
    params_default()); raises EpiError(INVALID) with the reader's message"""
    p = Params()
    C.pointer(p)[0] = base if base is not None else params_default()
    err = C.create_string_buffer(512)
    rc = load_library().epi_params_load(os.fsencode(path), int(bool(require_all)), C.byref(p), err, 512)
    if rc:
        raise EpiError(rc, err.value.decode())
    return p


def params_save(path, p):
    """one key=value line per key (Params::generate_paramfile, Params.cpp:98-110)"""
    rc = load_library().epi_params_save(os.fsencode(path), C.byref(p))
    if rc:
        raise EpiError(rc, f"cannot write {path}")


class GenCells(C.Structure):
    _fields_ = [("n", C.c_int32), ("nmemb", C.c_int32), ("nder", C.c_int32), ("nmx", C.c_int32),
                ("nmy", C.c_int32), ("state", _u8p), ("fix_origin", _ip), ("x", _dp), ("y", _dp), ("z", _dp),
                ("radius", _dp)]


def init_gen(params=None, nfix=16, der=8):
    """init_gen(nfix, der) of epi_intg (epi_intg/calc/init_gen.cpp:23-106)
    under `params` (default: epi5_MT's values): deck columns of the sheet,
    the dermis layers and the relaxed FIX ring, plus the sheet's (nmx, nmy)"""
    L = load_library()
    p = params if params is not None else params_default()
    g = GenCells()
    err = C.create_string_buffer(512)
    rc = L.epi_init_gen(C.byref(p), nfix, der, 0, C.byref(g), err, 512)
    if rc not in (0, 1):
        raise EpiError(rc, err.value.decode())
    n = g.n
    cols = {"state": np.zeros(n, np.uint8), "fix_origin": np.zeros(n, np.int32)}
    for k in ("x", "y", "z", "radius"):
        cols[k] = np.zeros(n)
    g = GenCells(0, 0, 0, 0, 0, _ptr(cols["state"], _u8p), _ptr(cols["fix_origin"], _ip), _ptr(cols["x"], _dp),
                 _ptr(cols["y"], _dp), _ptr(cols["z"], _dp), _ptr(cols["radius"], _dp))
    rc = L.epi_init_gen(C.byref(p), nfix, der, n, C.byref(g), err, 512)
    if rc:
        raise EpiError(rc, err.value.decode())
    return cols, (g.nmx, g.nmy)


def sheet_size(params=None):
    """(MEMB_NUM_X, MEMB_NUM_Y) of CalcParams::calculate_param (CalcParams.cpp:20-26)"""
    p = params if params is not None else params_default()
    x, y = C.c_int32(), C.c_int32()
    rc = load_library().epi_sheet_size(C.byref(p), C.byref(x), C.byref(y))
    if rc:
        raise EpiError(rc, "sheet size")
    return x.value, y.value


class RdParams(C.Structure):
    _fields_ = [("du", C.c_double), ("dp", C.c_double), ("gj0", C.c_double), ("kpa", C.POINTER(C.c_double))]


class MechParams(C.Structure):
    _fields_ = [("eps", C.c_double), ("k_adh", C.c_double), ("k_spring", C.c_double), ("k_bend", C.c_double)]


class RdGrad(C.Structure):
    _fields_ = [("du", C.c_double), ("dp", C.c_double), ("gj0", C.c_double), ("kpa", C.POINTER(C.c_double))]


_lib = None


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python epi_amd/build.py` "
                          "(or __graft_entry__.build()) first; there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.epi_create.argtypes = [C.c_int32, C.POINTER(Geometry), C.POINTER(vp)]
    L.epi_destroy.argtypes = [vp]
    L.epi_destroy.restype = None
    L.epi_last_error.argtypes = [vp]
    L.epi_last_error.restype = C.c_char_p
    L.epi_num_tissues.argtypes = [vp]
    L.epi_load_cells.argtypes = [vp, C.c_int32, C.POINTER(CellInput)]
    L.epi_value_init.argtypes = [vp, C.c_int32]
    L.epi_set_cell_values.argtypes = [vp, C.c_int32, _dp, _dp, _dp, _dp, _dp]
    L.epi_set_fields.argtypes = [vp, C.c_int32, _dp, _dp, _dp, _dp]
    L.epi_set_rng.argtypes = [vp, C.c_int32, _dp, C.c_int64]
    L.epi_rng_mt19937_64.argtypes = [C.c_uint64, _dp, C.c_int64]
    for f in ("epi_cell_interaction", "epi_pos_copy", "epi_cell_state_renew",
              "epi_cell_pos_periodic_fix", "epi_connect_cell", "epi_setup_map_lat",
              "epi_ca_calc_condition_reset", "epi_synchronize"):
        getattr(L, f).argtypes = [vp]
    L.epi_calc_zzmax.argtypes = [vp, _dp]
    for f in ("epi_calc_ext_stim", "epi_initialize_sc", "epi_calc_ca2p"):
        getattr(L, f).argtypes = [vp, _dp]
    L.epi_should_calc_ca.argtypes = [vp, _ip]
    L.epi_ca_substep.argtypes = [vp, C.c_int32]
    L.epi_step.argtypes = [vp, C.c_int64, C.POINTER(Driver)]
    L.epi_get_status.argtypes = [vp, C.c_int32, C.POINTER(Status)]
    L.epi_get_cells.argtypes = [vp, C.c_int32, C.POINTER(Cells)]
    L.epi_get_fields.argtypes = [vp, C.c_int32, _dp, _dp, _dp, _dp, _ip, _u8p]
    L.epi_set_positions.argtypes = [vp, C.c_int32, _dp, _dp, _dp]
    L.epi_set_state.argtypes = [vp, C.c_int32, _u8p, _dp]
    L.epi_set_lists.argtypes = [vp, C.c_int32, _ip, _ip, _dp]
    L.epi_set_maps.argtypes = [vp, C.c_int32, _ip, _u8p, _ip]
    L.epi_profile_enable.argtypes = [vp, C.c_uint64]
    L.epi_profile_sample.argtypes = [vp, C.c_int32, C.c_int32]
    L.epi_profile_launches.argtypes = [vp, C.c_int32, _i64p]
    L.epi_profile_read.argtypes = [vp, C.c_int32, C.POINTER(C.c_char_p), _dp, _i64p]
    L.epi_profile_reset.argtypes = [vp]
    L.epi_profile_kernel_count.argtypes = []
    L.epi_set_tuning.argtypes = [vp, C.c_int32, C.c_int32]
    L.epi_calc_ca2p_n.argtypes = [vp, C.POINTER(C.c_double), C.c_int32]
    L.epi_rd_begin.argtypes = [vp, C.POINTER(C.c_double)]
    L.epi_rd_forward.argtypes = [vp, C.POINTER(RdParams), C.c_int32]
    L.epi_rd_get_ca2p.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.epi_rd_adjoint.argtypes = [vp, C.POINTER(RdParams), C.c_int32, C.c_int32,
                                 C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_double), C.POINTER(RdGrad)]
    L.epi_rd_end.argtypes = [vp]
    L.epi_mech_begin.argtypes = [vp]
    L.epi_mech_forward.argtypes = [vp, C.POINTER(MechParams), C.c_int32]
    L.epi_mech_get_positions.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.epi_mech_adjoint.argtypes = [vp, C.POINTER(MechParams), C.c_int32, C.c_int32,
                                   C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_double), C.POINTER(MechParams)]
    L.epi_mech_end.argtypes = [vp]
    L.epi_get_tuning.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.epi_set_timestep.argtypes = [vp, C.c_int32, C.c_int64]
    L.epi_params_default.argtypes = [C.POINTER(Params)]
    L.epi_params_default.restype = None
    L.epi_params_load.argtypes = [C.c_char_p, C.c_int32, C.POINTER(Params), C.c_char_p, C.c_int32]
    L.epi_params_save.argtypes = [C.c_char_p, C.POINTER(Params)]
    L.epi_set_params.argtypes = [vp, C.c_int32, C.POINTER(Params)]
    L.epi_get_params.argtypes = [vp, C.c_int32, C.POINTER(Params)]
    L.epi_sheet_size.argtypes = [C.POINTER(Params), _ip, _ip]
    L.epi_init_gen.argtypes = [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32, C.POINTER(GenCells), C.c_char_p,
                               C.c_int32]
    L.epi_take_cell_stats.argtypes = [vp, C.c_int32, _i64p, C.c_int32, C.POINTER(C.c_int32)]
    L.epi_snapshot_begin.argtypes = [vp, C.c_int32, C.c_int32]
    L.epi_snapshot_wait.argtypes = [vp, C.c_int32, C.POINTER(Snapshot)]
    _lib = L
    return L


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def mt19937_64_uniform(seed, n):
    """std::mt19937_64(seed) through uniform_real_distribution<double>(0,1)."""
    out = np.empty(n, dtype=np.float64)
    rc = load_library().epi_rng_mt19937_64(C.c_uint64(seed), _ptr(out, _dp), n)
    if rc:
        raise EpiError(rc, "epi_rng_mt19937_64")
    return out


def _geo(g, row_stride=0, bucket_cap=0):
    return Geometry(g.nmx, g.nmy, g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.max_cells,
                    row_stride, bucket_cap)


class TissueSet:
    """One context: ``len(geometries)`` tissues advanced together."""

    def __init__(self, geometries, row_stride=0, bucket_cap=0):
        self.L = load_library()
        self.geos = list(geometries)
        self.T = len(self.geos)
        arr = (Geometry * self.T)(*[_geo(g, row_stride, bucket_cap) for g in self.geos])
        h = C.c_void_p()
        rc = self.L.epi_create(self.T, arr, C.byref(h))
        if rc:
            raise EpiError(rc, "epi_create")
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.epi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what=""):
        if rc:
            raise EpiError(rc, f"{what}: {self.L.epi_last_error(self.h).decode()}")

    # --- load (cman_init) ---
    def load(self, t, spec, value_init=False):
        c = spec["cells"]
        n = c["x"].size
        inp = CellInput()
        inp.n, inp.nmemb, inp.nder = n, spec["nmemb"], spec["nder"]
        conv = {}
        for k, dt, pt in (("state", np.uint8, _u8p), ("fix_origin", np.int32, _ip),
                          ("x", np.float64, _dp), ("y", np.float64, _dp), ("z", np.float64, _dp),
                          ("radius", np.float64, _dp), ("ca2p", np.float64, _dp),
                          ("ca2p_avg", np.float64, _dp), ("ip3", np.float64, _dp),
                          ("ex_inert", np.float64, _dp), ("agek", np.float64, _dp),
                          ("ageb", np.float64, _dp), ("ex_fat", np.float64, _dp),
                          ("in_fat", np.float64, _dp), ("spr_nat_len", np.float64, _dp),
                          ("rest_div_times", np.int32, _ip), ("is_malignant", np.uint8, _u8p),
                          ("pair", np.int32, _ip)):
            conv[k] = np.ascontiguousarray(c[k], dtype=dt)
            setattr(inp, k, _ptr(conv[k], pt))
        self._keep[("in", t)] = conv
        self._chk(self.L.epi_load_cells(self.h, t, C.byref(inp)), "load_cells")
        self._chk(self.L.epi_connect_cell(self.h), "connect_cell")
        vals = spec.get("values")
        if value_init or vals is None:
            self._chk(self.L.epi_value_init(self.h, t), "value_init")
        else:
            self.set_cell_values(t, **vals)
        f = spec.get("fields")
        if f is not None:
            self.set_fields(t, f["atp"], f["atp"], f["b"], f["b"])

    def set_cell_values(self, t, ca2p=None, ip3=None, ex_inert=None, ca2p_avg=None, gj=None):
        cv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        ca2p, ip3, ex_inert, ca2p_avg, gj = map(cv, (ca2p, ip3, ex_inert, ca2p_avg, gj))
        self._chk(self.L.epi_set_cell_values(self.h, t, _ptr(ca2p, _dp), _ptr(ip3, _dp),
                                             _ptr(ex_inert, _dp), _ptr(ca2p_avg, _dp),
                                             _ptr(gj, _dp)), "set_cell_values")

    def set_fields(self, t, atp=None, atp_next=None, b=None, b_next=None):
        cv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
        a, an, bb, bn = map(cv, (atp, atp_next, b, b_next))
        self._chk(self.L.epi_set_fields(self.h, t, _ptr(a, _dp), _ptr(an, _dp), _ptr(bb, _dp),
                                        _ptr(bn, _dp)), "set_fields")

    def set_rng(self, t, draws):
        d = np.ascontiguousarray(draws, dtype=np.float64)
        self._chk(self.L.epi_set_rng(self.h, t, _ptr(d, _dp), d.size), "set_rng")

    # --- reference call surface (every tissue) ---
    def cell_interaction(self): self._chk(self.L.epi_cell_interaction(self.h), "cell_interaction")
    def pos_copy(self): self._chk(self.L.epi_pos_copy(self.h), "pos_copy")
    def cell_state_renew(self): self._chk(self.L.epi_cell_state_renew(self.h), "cell_state_renew")
    def cell_pos_periodic_fix(self): self._chk(self.L.epi_cell_pos_periodic_fix(self.h), "periodic_fix")
    def connect_cell(self): self._chk(self.L.epi_connect_cell(self.h), "connect_cell")
    def setup_map_lat(self): self._chk(self.L.epi_setup_map_lat(self.h), "setup_map_lat")
    def ca_calc_condition_reset(self): self._chk(self.L.epi_ca_calc_condition_reset(self.h), "reset")
    def synchronize(self): self._chk(self.L.epi_synchronize(self.h), "synchronize")

    def _zz(self, zzmax):
        if zzmax is None:
            return None
        z = np.ascontiguousarray(np.broadcast_to(np.asarray(zzmax, dtype=np.float64), (self.T,)))
        self._keep["zz"] = z
        return _ptr(z, _dp)

    def calc_zzmax(self):
        out = np.zeros(self.T)
        self._chk(self.L.epi_calc_zzmax(self.h, _ptr(out, _dp)), "calc_zzmax")
        return out

    def calc_ext_stim(self, zzmax): self._chk(self.L.epi_calc_ext_stim(self.h, self._zz(zzmax)), "ext_stim")
    def initialize_sc(self, zzmax): self._chk(self.L.epi_initialize_sc(self.h, self._zz(zzmax)), "initialize_sc")
    def calc_ca2p(self, zzmax): self._chk(self.L.epi_calc_ca2p(self.h, self._zz(zzmax)), "calc_ca2p")
    def ca_substep(self, iz_bound): self._chk(self.L.epi_ca_substep(self.h, iz_bound), "ca_substep")
    def calc_ca2p_n(self, zzmax, n_sub):
        self._chk(self.L.epi_calc_ca2p_n(self.h, self._zz(zzmax), int(n_sub)), "calc_ca2p_n")

    def should_calc_ca(self):
        f = np.zeros(self.T, dtype=np.int32)
        self._chk(self.L.epi_should_calc_ca(self.h, _ptr(f, _ip)), "should_calc_ca")
        return f.astype(bool)

    def step(self, drivers, nsteps=1):
        arr = (Driver * self.T)(*drivers)
        rc = self.L.epi_step(self.h, nsteps, arr)
        for i in range(self.T):
            drivers[i].step, drivers[i].force_cornif = arr[i].step, arr[i].force_cornif
            drivers[i].num_sc = arr[i].num_sc
        self._chk(rc, "step")

    # --- kernel timing ---
    def profile_enable(self, names=None):
        """Time the named kernels (all when names is None) with HIP events."""
        n = self.L.epi_profile_kernel_count()
        mask = 0
        for k in range(n):
            if names is None or self.kernel_name(k) in names:
                mask |= 1 << k
        self._chk(self.L.epi_profile_enable(self.h, mask), "profile_enable")

    def profile_sample(self, period=1, run=2):
        """period 1: every launch of an enabled kernel is timed.  period > 1:
        runs of `run` consecutive launches carry events, one run every
        `period` launches; the first launch of a run is not counted."""
        self._chk(self.L.epi_profile_sample(self.h, int(period), int(run)), "profile_sample")

    def kernel_name(self, k):
        nm = C.c_char_p()
        self._chk(self.L.epi_profile_read(self.h, k, C.byref(nm), None, None), "profile_read")
        return nm.value.decode()

    def profile(self):
        """{kernel: (total_ms, launches)} for kernels with launches."""
        out = {}
        for k in range(self.L.epi_profile_kernel_count()):
            nm, ms, cnt = C.c_char_p(), C.c_double(), C.c_int64()
            self._chk(self.L.epi_profile_read(self.h, k, C.byref(nm), C.byref(ms), C.byref(cnt)), "profile_read")
            if cnt.value:
                out[nm.value.decode()] = (ms.value, cnt.value)
        return out

    def profile_launches(self):
        """{kernel: launches} of the enabled kernels since the last reset,
        timed or not (with sampling only some launches are timed)."""
        out = {}
        for k in range(self.L.epi_profile_kernel_count()):
            n = C.c_int64()
            self._chk(self.L.epi_profile_launches(self.h, k, C.byref(n)), "profile_launches")
            if n.value:
                out[self.kernel_name(k)] = n.value
        return out

    def profile_reset(self):
        self._chk(self.L.epi_profile_reset(self.h), "profile_reset")

    # --- frozen mechanics window ---
    def _mech_params(self, params):
        arr = (MechParams * self.T)()
        for t, p in enumerate(params):
            q = dict(MECH_DEFAULT)
            q.update(p)
            arr[t] = MechParams(q["eps"], q["k_adh"], q["k_spring"], q["k_bend"])
        return arr

    def mech_begin(self):
        self._chk(self.L.epi_mech_begin(self.h), "mech_begin")

    def mech_forward(self, params, n_steps):
        self._chk(self.L.epi_mech_forward(self.h, self._mech_params(params), int(n_steps)), "mech_forward")

    def mech_get_positions(self, t):
        n = self.status(t)[0].n
        out = np.zeros(3 * n)
        self._chk(self.L.epi_mech_get_positions(self.h, t, _ptr(out, _dp)), "mech_get_positions")
        return out.reshape(3, n)

    def mech_adjoint(self, params, n_steps, checkpoint_every, targets):
        tg = [np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1)) for x in targets]
        tp = (C.POINTER(C.c_double) * self.T)(*[_ptr(x, _dp) for x in tg])
        loss = np.zeros(self.T)
        grads = (MechParams * self.T)()
        self._chk(self.L.epi_mech_adjoint(self.h, self._mech_params(params), int(n_steps), int(checkpoint_every), tp,
                                          _ptr(loss, _dp), grads), "mech_adjoint")
        return loss, [{"eps": g.eps, "k_adh": g.k_adh, "k_spring": g.k_spring, "k_bend": g.k_bend} for g in grads]

    def mech_end(self):
        self._chk(self.L.epi_mech_end(self.h), "mech_end")

    # --- frozen reaction-diffusion window ---
    def _rd_params(self, params):
        arr = (RdParams * self.T)()
        keep = []
        for t, p in enumerate(params):
            k = p.get("kpa")
            if k is not None:
                k = np.ascontiguousarray(k, dtype=np.float64)
                keep.append(k)
            arr[t] = RdParams(p.get("du", 0.01), p.get("dp", 0.1), p.get("gj0", 0.99),
                              None if k is None else _ptr(k, _dp))
        return arr, keep

    def rd_begin(self, zzmax):
        self._chk(self.L.epi_rd_begin(self.h, self._zz(zzmax)), "rd_begin")

    def rd_forward(self, params, n_sub):
        arr, keep = self._rd_params(params)
        self._chk(self.L.epi_rd_forward(self.h, arr, int(n_sub)), "rd_forward")

    def rd_get_ca2p(self, t):
        out = np.zeros(self.status(t)[0].n)
        self._chk(self.L.epi_rd_get_ca2p(self.h, t, _ptr(out, _dp)), "rd_get_ca2p")
        return out

    def rd_adjoint(self, params, n_sub, checkpoint_every, targets):
        arr, keep = self._rd_params(params)
        tg = [np.ascontiguousarray(x, dtype=np.float64) for x in targets]
        tp = (C.POINTER(C.c_double) * self.T)(*[_ptr(x, _dp) for x in tg])
        loss = np.zeros(self.T)
        gk = [np.zeros(self.status(t)[0].n) for t in range(self.T)]
        grads = (RdGrad * self.T)(*[RdGrad(0, 0, 0, _ptr(g, _dp)) for g in gk])
        self._chk(self.L.epi_rd_adjoint(self.h, arr, int(n_sub), int(checkpoint_every), tp, _ptr(loss, _dp), grads),
                  "rd_adjoint")
        return loss, [{"du": grads[t].du, "dp": grads[t].dp, "gj0": grads[t].gj0, "kpa": gk[t]}
                      for t in range(self.T)]

    def rd_end(self):
        self._chk(self.L.epi_rd_end(self.h), "rd_end")

    # --- launch shapes ---
    def set_tuning(self, tuning):
        """Apply a {key: value} dict of launch shapes (keys of TUNING_KEYS)."""
        for k, v in tuning.items():
            self._chk(self.L.epi_set_tuning(self.h, TUNING_KEYS[k], int(v)), f"set_tuning {k}={v}")

    def get_tuning(self):
        out = {}
        for k, key in TUNING_KEYS.items():
            v = C.c_int32()
            self._chk(self.L.epi_get_tuning(self.h, key, C.byref(v)), "get_tuning")
            out[k] = v.value
        return out

    # --- asynchronous snapshots ---
    def snapshot_begin(self, t, slot=0):
        """queue tissue t's output record into snapshot slot `slot` behind
        the work queued so far; returns at once"""
        self._chk(self.L.epi_snapshot_begin(self.h, int(t), int(slot)), "snapshot_begin")

    def snapshot_wait(self, slot=0):
        """wait for slot's download; the record as numpy copies"""
        o = Snapshot()
        self._chk(self.L.epi_snapshot_wait(self.h, int(slot), C.byref(o)), "snapshot_wait")
        n = o.n
        out = {"tissue": o.tissue, "n": n}
        for k, _ in Snapshot._fields_[2:]:
            out[k] = np.ctypeslib.as_array(getattr(o, k), shape=(n,)).copy() if n else np.zeros(0)
        return out

    # --- state exchange ---
    def status(self, t):
        s = Status()
        rc = self.L.epi_get_status(self.h, t, C.byref(s))
        return s, rc

    def get_cells(self, t):
        s, _ = self.status(t)
        n = s.n
        ptrs = np.zeros(n + 1, dtype=np.int32)
        probe = Cells()
        probe.conn_ptr = _ptr(ptrs, _ip)
        self._chk(self.L.epi_get_cells(self.h, t, C.byref(probe)), "get_cells")
        nnz = int(ptrs[n])
        out = {k: np.zeros(n, dtype=dt) for k, dt in CELL_DTYPES.items()}
        out["lat"] = np.zeros(3 * n, dtype=np.int32)
        out["conn_ptr"] = np.zeros(n + 1, dtype=np.int32)
        out["conn"] = np.zeros(max(nnz, 1), dtype=np.int32)
        out["gj"] = np.zeros(max(nnz, 1), dtype=np.float64)
        out["memb_adj"] = np.zeros(6 * probe.nmemb, dtype=np.int32)
        out["cst"] = np.zeros((n, 3), dtype=np.int64)
        cs = Cells()
        for k, dt in CELL_DTYPES.items():
            pt = {np.uint8: _u8p, np.int32: _ip, np.float64: _dp, np.int64: _i64p}[dt]
            setattr(cs, k, _ptr(out[k], pt))
        cs.lat = _ptr(out["lat"], _ip)
        cs.conn_ptr = _ptr(out["conn_ptr"], _ip)
        cs.conn = _ptr(out["conn"], _ip)
        cs.gj = _ptr(out["gj"], _dp)
        cs.conn_cap = out["conn"].size
        cs.memb_adj = _ptr(out["memb_adj"], _ip)
        cs.cst = _ptr(out["cst"], _i64p)
        self._chk(self.L.epi_get_cells(self.h, t, C.byref(cs)), "get_cells")
        out["conn"] = out["conn"][:nnz]
        out["gj"] = out["gj"][:nnz]
        out["n"], out["nmemb"], out["nder"], out["sw"] = cs.n, cs.nmemb, cs.nder, cs.sw
        return out

    # --- runtime parameters (epi_intg/calc/CalcParams.cpp:16-279) ---
    def set_params(self, t, p):
        """tissue t's parameters, read from the next call on (a Params, or a
        dict of keys changed from epi5_MT's values)"""
        if isinstance(p, dict):
            p = self.get_params(t).replace(**p)
        self._chk(self.L.epi_set_params(self.h, t, C.byref(p)), "set_params")

    def get_params(self, t):
        p = Params()
        self._chk(self.L.epi_get_params(self.h, t, C.byref(p)), "get_params")
        return p

    # --- cell_stat (cellmanager.cpp:338-352, 559-567) ---
    def set_timestep(self, t, step):
        self._chk(self.L.epi_set_timestep(self.h, t, int(step)), "set_timestep")

    def take_cell_stats(self, t):
        """the records (k_aging_start, k_cornified, k_disap) of the DEAD
        cells removed since the last call, in removal order"""
        k = C.c_int32(0)
        cap = self.geos[t].max_cells
        out = np.zeros((cap, 3), dtype=np.int64)
        self._chk(self.L.epi_take_cell_stats(self.h, t, _ptr(out, _i64p), cap, C.byref(k)), "take_cell_stats")
        return out[:k.value].copy()

    def get_fields(self, t):
        g = self.geos[t]
        shape = (g.nx + 1, g.ny + 1, g.nz + 1)
        sz = int(np.prod(shape))
        o = {k: np.zeros(sz, dtype=np.float64) for k in ("atp", "atp_next", "b", "b_next")}
        o["cmap1"] = np.zeros(sz, dtype=np.int32)
        o["cmap2"] = np.zeros(sz, dtype=np.uint8)
        self._chk(self.L.epi_get_fields(self.h, t, _ptr(o["atp"], _dp), _ptr(o["atp_next"], _dp),
                                        _ptr(o["b"], _dp), _ptr(o["b_next"], _dp),
                                        _ptr(o["cmap1"], _ip), _ptr(o["cmap2"], _u8p)), "get_fields")
        return {k: v.reshape(shape) for k, v in o.items()}

    def set_positions(self, t, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        self._chk(self.L.epi_set_positions(self.h, t, _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp)), "set_positions")

    def set_state(self, t, state=None, agek=None):
        s = None if state is None else np.ascontiguousarray(state, dtype=np.uint8)
        a = None if agek is None else np.ascontiguousarray(agek, dtype=np.float64)
        self._chk(self.L.epi_set_state(self.h, t, _ptr(s, _u8p), _ptr(a, _dp)), "set_state")

    def set_lists(self, t, conn_ptr, conn, gj=None):
        cp = np.ascontiguousarray(conn_ptr, dtype=np.int32)
        cn = np.ascontiguousarray(conn, dtype=np.int32)
        g = None if gj is None else np.ascontiguousarray(gj, dtype=np.float64)
        self._chk(self.L.epi_set_lists(self.h, t, _ptr(cp, _ip), _ptr(cn, _ip), _ptr(g, _dp)), "set_lists")

    def set_maps(self, t, cmap1=None, cmap2=None, lat=None):
        c1 = None if cmap1 is None else np.ascontiguousarray(cmap1, dtype=np.int32).ravel()
        c2 = None if cmap2 is None else np.ascontiguousarray(cmap2, dtype=np.uint8).ravel()
        la = None if lat is None else np.ascontiguousarray(lat, dtype=np.int32).ravel()
        self._chk(self.L.epi_set_maps(self.h, t, _ptr(c1, _ip), _ptr(c2, _u8p), _ptr(la, _ip)), "set_maps")


class Tissue:
    """Single-tissue view with the same methods as the oracle wrapper."""

    def __init__(self, geometry, **kw):
        self.set = TissueSet([geometry], **kw)
        self.geo = geometry

    def load(self, spec, value_init=False): self.set.load(0, spec, value_init)
    def set_rng(self, draws): self.set.set_rng(0, draws)
    def set_cell_values(self, **kw): self.set.set_cell_values(0, **kw)
    def set_fields(self, *a, **kw): self.set.set_fields(0, *a, **kw)
    def cell_interaction(self): self.set.cell_interaction()
    def pos_copy(self): self.set.pos_copy()
    def cell_state_renew(self): self.set.cell_state_renew()
    def cell_pos_periodic_fix(self): self.set.cell_pos_periodic_fix()
    def connect_cell(self): self.set.connect_cell()
    def calc_zzmax(self): return float(self.set.calc_zzmax()[0])
    def setup_map_lat(self): self.set.setup_map_lat()
    def calc_ext_stim(self, zzmax): self.set.calc_ext_stim(zzmax)
    def initialize_sc(self, zzmax): self.set.initialize_sc(zzmax)
    def calc_ca2p(self, zzmax): self.set.calc_ca2p(zzmax)
    def ca_substep(self, iz_bound): self.set.ca_substep(iz_bound)
    def should_calc_ca(self): return bool(self.set.should_calc_ca()[0])
    def ca_calc_condition_reset(self): self.set.ca_calc_condition_reset()
    def step(self, driver, nsteps=1): self.set.step([driver], nsteps)
    def status(self): return self.set.status(0)[0]
    def rng_cursor(self): return int(self.status().rng_pos)
    def error(self): return int(self.set.status(0)[0].err)
    def get_cells(self): return self.set.get_cells(0)
    def set_timestep(self, step): self.set.set_timestep(0, step)
    def take_cell_stats(self): return self.set.take_cell_stats(0)
    def set_params(self, p): self.set.set_params(0, p)
    def get_params(self): return self.set.get_params(0)
    def get_fields(self): return self.set.get_fields(0)
    def set_positions(self, x, y, z): self.set.set_positions(0, x, y, z)
    def set_state(self, state=None, agek=None): self.set.set_state(0, state, agek)
    def set_lists(self, conn_ptr, conn, gj=None): self.set.set_lists(0, conn_ptr, conn, gj)
    def set_maps(self, cmap1=None, cmap2=None, lat=None): self.set.set_maps(0, cmap1, cmap2, lat)
    def close(self): self.set.close()
