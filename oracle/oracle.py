"""ctypes wrapper of the CPU restatement (oracle/libepi_oracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libepi_oracle.so")
# the same restatement with OpenMP threads over the loops epi5_MT runs in
# parallel: the CPU throughput baseline only (its list and accumulation
# orders follow the thread schedule, like the reference's under TBB)
LIB_PATH_MT = os.path.join(HERE, "libepi_oracle_mt.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


class Geo(C.Structure):
    _fields_ = [("nmx", C.c_int32), ("nmy", C.c_int32), ("nx", C.c_int32),
                ("ny", C.c_int32), ("nz", C.c_int32), ("lx", C.c_double),
                ("ly", C.c_double), ("lz", C.c_double), ("max_cells", C.c_int32)]


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
                ("memb_adj", _ip)]


class Driver(C.Structure):
    _fields_ = [("step", C.c_int64), ("force_cornif", C.c_int32), ("num_sc", C.c_int32),
                ("ca_every", C.c_int32), ("clean_every", C.c_int32)]


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


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


HEADER = os.path.join(os.path.dirname(HERE), "include", "epi_amd.h")


def header_param_fields(path=HEADER):
    """the fields of `typedef struct epi_params` as the C header declares
    them (the oracle's C code includes the same header): (name, ctypes type)"""
    text = open(path).read()
    a = text.index("typedef struct epi_params {") + len("typedef struct epi_params {")
    body = text[a:text.index("} epi_params;")]
    kinds = {"int32_t": C.c_int32, "double": C.c_double}
    out = []
    for stmt in body.split(";"):
        stmt = " ".join(stmt.split())
        if not stmt:
            continue
        ty, names = stmt.split(" ", 1)
        for nm in names.split(","):
            nm = nm.strip()
            if "[" in nm:
                base, n = nm.split("[")
                out.append((base, C.c_char * int(n.rstrip("]"))))
            else:
                out.append((nm, kinds[ty]))
    return out


class Params(C.Structure):
    _fields_ = header_param_fields()

    def to_dict(self):
        return {k: (getattr(self, k).decode() if k == "outputdir" else getattr(self, k)) for k, _ in self._fields_}


PARAM_KEYS = [k for k, _ in Params._fields_]


class GenCells(C.Structure):
    _fields_ = [("n", C.c_int32), ("nmemb", C.c_int32), ("nder", C.c_int32), ("nmx", C.c_int32),
                ("nmy", C.c_int32), ("state", _u8p), ("fix_origin", _ip), ("x", _dp), ("y", _dp), ("z", _dp),
                ("radius", _dp)]


def init_gen(params, nfix=16, der=8):
    """the oracle's init_gen (epi_intg/calc/init_gen.cpp:23-106): deck
    columns and the sheet's (nmx, nmy)"""
    L = lib()
    p = Params()
    C.memmove(C.byref(p), C.byref(params), C.sizeof(Params))
    g = GenCells()
    rc = L.eo_init_gen(C.byref(p), nfix, der, 0, C.byref(g))
    n = g.n
    cols = {"state": np.zeros(n, np.uint8), "fix_origin": np.zeros(n, np.int32)}
    for k in ("x", "y", "z", "radius"):
        cols[k] = np.zeros(n)
    g = GenCells(0, 0, 0, 0, 0, cols["state"].ctypes.data_as(_u8p), cols["fix_origin"].ctypes.data_as(_ip),
                 cols["x"].ctypes.data_as(_dp), cols["y"].ctypes.data_as(_dp), cols["z"].ctypes.data_as(_dp),
                 cols["radius"].ctypes.data_as(_dp))
    rc = L.eo_init_gen(C.byref(p), nfix, der, n, C.byref(g))
    if rc:
        raise RuntimeError(f"eo_init_gen: {rc}")
    return cols, (g.nmx, g.nmy)


_libs = {}


class RdParams(C.Structure):
    _fields_ = [("du", C.c_double), ("dp", C.c_double), ("gj0", C.c_double), ("kpa", _dp)]


class MechParams(C.Structure):
    _fields_ = [("eps", C.c_double), ("k_adh", C.c_double), ("k_spring", C.c_double), ("k_bend", C.c_double)]


MECH_DEFAULT = {"eps": 0.01, "k_adh": 3.0, "k_spring": 25.0, "k_bend": 0.5}


class RdGrad(C.Structure):
    _fields_ = [("du", C.c_double), ("dp", C.c_double), ("gj0", C.c_double), ("kpa", _dp)]


def params_default():
    """epi5_MT's values as the oracle restates them"""
    p = Params()
    lib().eo_params_default(C.byref(p))
    return p


def lib(mt=False):
    path = LIB_PATH_MT if mt else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.eo_create.restype = C.c_void_p
        L.eo_params_default.argtypes = [C.POINTER(Params)]
        L.eo_params_default.restype = None
        L.eo_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
        L.eo_set_params.restype = None
        L.eo_get_params.argtypes = [C.c_void_p, C.POINTER(Params)]
        L.eo_get_params.restype = None
        L.eo_init_gen.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.eo_create.argtypes = [C.POINTER(Geo)]
        L.eo_destroy.argtypes = [C.c_void_p]
        L.eo_error.argtypes = [C.c_void_p]
        L.eo_clear_error.argtypes = [C.c_void_p]
        L.eo_create_cell.argtypes = [C.c_void_p, C.c_int, C.c_int] + [C.c_double] * 13 + [C.c_int, C.c_int]
        L.eo_link_pair.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.eo_finish_load.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.eo_value_init.argtypes = [C.c_void_p]
        L.eo_set_rng.argtypes = [C.c_void_p, _dp, C.c_int64]
        L.eo_rng_cursor.argtypes = [C.c_void_p]
        L.eo_rng_cursor.restype = C.c_int64
        L.eo_mt19937_64_uniform.argtypes = [C.c_uint64, _dp, C.c_int64]
        L.eo_mt19937_64_raw.argtypes = [C.c_uint64, C.POINTER(C.c_uint64), C.c_int64]
        for f in ("eo_cell_interaction", "eo_pos_copy", "eo_cell_state_renew",
                  "eo_cell_pos_periodic_fix", "eo_connect_cell", "eo_setup_map_lat",
                  "eo_ca_calc_condition_reset", "eo_clean_up"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.eo_calc_zzmax.argtypes = [C.c_void_p]
        L.eo_calc_zzmax.restype = C.c_double
        for f in ("eo_calc_ext_stim", "eo_initialize_sc", "eo_calc_ca2p"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_double]
        L.eo_ca_substep.argtypes = [C.c_void_p, C.c_int]
        L.eo_calc_ca2p_n.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.eo_should_calc_ca.argtypes = [C.c_void_p]
        L.eo_step.argtypes = [C.c_void_p, C.POINTER(Driver)]
        L.eo_mech_forward.argtypes = [C.c_void_p, C.POINTER(MechParams), C.c_int, _dp]
        L.eo_mech_adjoint.argtypes = [C.c_void_p, C.POINTER(MechParams), C.c_int, _dp, _dp, C.POINTER(MechParams)]
        L.eo_rd_forward.argtypes = [C.c_void_p, C.c_double, C.POINTER(RdParams), C.c_int, _dp]
        L.eo_rd_adjoint.argtypes = [C.c_void_p, C.c_double, C.POINTER(RdParams), C.c_int, _dp, _dp,
                                    C.POINTER(RdGrad)]
        L.eo_get_cells.argtypes = [C.c_void_p, C.POINTER(Cells)]
        L.eo_set_cell_values.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.eo_get_fields.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip, _u8p]
        L.eo_set_fields.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.eo_set_maps.argtypes = [C.c_void_p, _ip, _u8p, _ip]
        L.eo_set_positions.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.eo_set_state.argtypes = [C.c_void_p, _u8p, _dp]
        L.eo_set_lists.argtypes = [C.c_void_p, _ip, _ip, _dp]
        L.eo_p_diff_x.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.eo_p_diff_x.restype = C.c_double
        L.eo_coef.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.eo_coef.restype = C.c_double
        L.eo_tri_bend.argtypes = [C.c_void_p, C.c_int, _dp]
        L.eo_set_timestep.argtypes = [C.c_void_p, C.c_int64]
        L.eo_timestep.argtypes = [C.c_void_p]
        L.eo_timestep.restype = C.c_int64
        L.eo_get_cst.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.eo_take_cell_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int)]
        _libs[path] = L
    return _libs[path]


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def mt19937_64_uniform(seed, n):
    out = np.empty(n, dtype=np.float64)
    lib().eo_mt19937_64_uniform(C.c_uint64(seed), ptr(out, _dp), n)
    return out


def mt19937_64_raw(seed, n):
    out = np.empty(n, dtype=np.uint64)
    lib().eo_mt19937_64_raw(C.c_uint64(seed), out.ctypes.data_as(C.POINTER(C.c_uint64)), n)
    return out


class OracleTissue:
    """One tissue in the CPU restatement, with the reference call surface."""

    def __init__(self, geometry, mt=False):
        self.L = lib(mt)
        g = geometry
        self.geo = g
        self._g = Geo(g.nmx, g.nmy, g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.max_cells)
        self.h = self.L.eo_create(C.byref(self._g))
        self._rng = None
        self.shape = (g.nx + 1, g.ny + 1, g.nz + 1)

    def __del__(self):
        try:
            if self.h:
                self.L.eo_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # --- load (cman_init) ---
    def load(self, spec, value_init=False):
        c = spec["cells"]
        n = c["x"].size
        for i in range(n):
            self.L.eo_create_cell(self.h, int(c["state"][i]), int(c["fix_origin"][i]),
                                  float(c["x"][i]), float(c["y"][i]), float(c["z"][i]),
                                  float(c["radius"][i]), float(c["ca2p"][i]),
                                  float(c["ca2p_avg"][i]), float(c["ip3"][i]),
                                  float(c["ex_inert"][i]), float(c["agek"][i]),
                                  float(c["ageb"][i]), float(c["ex_fat"][i]),
                                  float(c["in_fat"][i]), float(c["spr_nat_len"][i]),
                                  int(c["rest_div_times"][i]), int(c["is_malignant"][i]))
        for i in range(n):
            p = int(c["pair"][i])
            if p > i:
                self.L.eo_link_pair(self.h, i, p)
        rc = self.L.eo_finish_load(self.h, spec["nmemb"], spec["nder"])
        if rc:
            raise RuntimeError(f"finish_load failed {rc}")
        self.connect_cell()
        vals = spec.get("values")
        if value_init or vals is None:
            self.L.eo_value_init(self.h)
        else:
            self.set_cell_values(**vals)
        f = spec.get("fields")
        if f is not None:
            self.set_fields(f["atp"], f["atp"], f["b"], f["b"])
        self.check()

    def set_cell_values(self, ca2p=None, ip3=None, ex_inert=None, ca2p_avg=None, gj=None):
        cv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        ca2p, ip3, ex_inert, ca2p_avg, gj = map(cv, (ca2p, ip3, ex_inert, ca2p_avg, gj))
        self.L.eo_set_cell_values(self.h, ptr(ca2p, _dp), ptr(ip3, _dp), ptr(ex_inert, _dp),
                                  ptr(ca2p_avg, _dp), ptr(gj, _dp))

    def set_rng(self, draws):
        self._rng = np.ascontiguousarray(draws, dtype=np.float64)
        self.L.eo_set_rng(self.h, ptr(self._rng, _dp), self._rng.size)

    def rng_cursor(self):
        return int(self.L.eo_rng_cursor(self.h))

    def error(self):
        return int(self.L.eo_error(self.h))

    def check(self):
        e = self.error()
        if e:
            raise RuntimeError(f"oracle error code {e}")

    # --- reference call surface ---
    def cell_interaction(self): self.L.eo_cell_interaction(self.h)
    def pos_copy(self): self.L.eo_pos_copy(self.h)
    def cell_state_renew(self): self.L.eo_cell_state_renew(self.h)
    def cell_pos_periodic_fix(self): self.L.eo_cell_pos_periodic_fix(self.h)
    def connect_cell(self): self.L.eo_connect_cell(self.h)
    def calc_zzmax(self): return float(self.L.eo_calc_zzmax(self.h))
    def setup_map_lat(self): self.L.eo_setup_map_lat(self.h)
    def calc_ext_stim(self, zzmax): self.L.eo_calc_ext_stim(self.h, zzmax)
    def initialize_sc(self, zzmax): self.L.eo_initialize_sc(self.h, zzmax)
    def calc_ca2p(self, zzmax): self.L.eo_calc_ca2p(self.h, zzmax)
    def ca_substep(self, iz_bound): self.L.eo_ca_substep(self.h, iz_bound)
    def calc_ca2p_n(self, zzmax, n_sub): self.L.eo_calc_ca2p_n(self.h, zzmax, n_sub)
    def should_calc_ca(self): return bool(self.L.eo_should_calc_ca(self.h))
    def ca_calc_condition_reset(self): self.L.eo_ca_calc_condition_reset(self.h)
    def clean_up(self): self.L.eo_clean_up(self.h)

    # --- cell_stat (cellmanager.cpp:338-352, 559-567) ---
    def set_timestep(self, step): self.L.eo_set_timestep(self.h, int(step))

    def get_params(self):
        p = Params()
        self.L.eo_get_params(self.h, C.byref(p))
        return p

    def set_params(self, p):
        """a Params, or a dict of keys changed from the current ones"""
        if isinstance(p, dict):
            q = self.get_params()
            for k, v in p.items():
                if k not in PARAM_KEYS:
                    raise KeyError(k)
                setattr(q, k, v.encode() if k == "outputdir" else v)
            p = q
        self.L.eo_set_params(self.h, C.byref(p))

    def timestep(self): return int(self.L.eo_timestep(self.h))

    def get_cst(self):
        """per cell (k_aging_start, k_cornified, k_disap), -1 = never set"""
        n = self.get_cells()["n"]
        out = np.empty((n, 3), dtype=np.int64)
        self.L.eo_get_cst(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out

    def take_cell_stats(self):
        """the records of the DEAD cells removed since the last call, in
        removal order; the store is cleared"""
        cap = self.geo.max_cells
        out = np.empty((cap, 3), dtype=np.int64)
        k = C.c_int(0)
        rc = self.L.eo_take_cell_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)), cap, C.byref(k))
        if rc:
            raise RuntimeError(f"take_cell_stats: {rc}")
        return out[:k.value].copy()

    def step(self, driver):
        return int(self.L.eo_step(self.h, C.byref(driver)))

    # --- frozen mechanics window ---
    def mech_forward(self, n_steps, **params):
        p = dict(MECH_DEFAULT)
        p.update(params)
        n = self.get_cells()["x"].size
        out = np.zeros(3 * n)
        rc = self.L.eo_mech_forward(self.h, C.byref(MechParams(**p)), n_steps, out.ctypes.data_as(_dp))
        if rc:
            raise RuntimeError(f"eo_mech_forward {rc}")
        return out.reshape(3, n)

    def mech_adjoint(self, n_steps, target, **params):
        p = dict(MECH_DEFAULT)
        p.update(params)
        tgt = np.ascontiguousarray(np.asarray(target, dtype=np.float64).reshape(-1))
        g = MechParams()
        loss = C.c_double()
        rc = self.L.eo_mech_adjoint(self.h, C.byref(MechParams(**p)), n_steps, tgt.ctypes.data_as(_dp),
                                    C.byref(loss), C.byref(g))
        if rc:
            raise RuntimeError(f"eo_mech_adjoint {rc}")
        return loss.value, {"eps": g.eps, "k_adh": g.k_adh, "k_spring": g.k_spring, "k_bend": g.k_bend}

    # --- frozen reaction-diffusion window ---
    def _rd_params(self, du, dp, gj0, kpa):
        k = None if kpa is None else np.ascontiguousarray(kpa, dtype=np.float64)
        p = RdParams(du, dp, gj0, None if k is None else k.ctypes.data_as(_dp))
        return p, k

    def rd_forward(self, zzmax, n_sub, du=0.01, dp=0.1, gj0=0.99, kpa=None):
        p, keep = self._rd_params(du, dp, gj0, kpa)
        n = self.get_cells()["x"].size
        out = np.zeros(n)
        rc = self.L.eo_rd_forward(self.h, zzmax, C.byref(p), n_sub, out.ctypes.data_as(_dp))
        if rc:
            raise RuntimeError(f"eo_rd_forward {rc}")
        return out

    def rd_adjoint(self, zzmax, n_sub, target, du=0.01, dp=0.1, gj0=0.99, kpa=None):
        p, keep = self._rd_params(du, dp, gj0, kpa)
        n = self.get_cells()["x"].size
        tgt = np.ascontiguousarray(target, dtype=np.float64)
        gk = np.zeros(n)
        g = RdGrad(0, 0, 0, gk.ctypes.data_as(_dp))
        loss = C.c_double()
        rc = self.L.eo_rd_adjoint(self.h, zzmax, C.byref(p), n_sub, tgt.ctypes.data_as(_dp), C.byref(loss),
                                  C.byref(g))
        if rc:
            raise RuntimeError(f"eo_rd_adjoint {rc}")
        return loss.value, {"du": g.du, "dp": g.dp, "gj0": g.gj0, "kpa": gk}

    # --- state exchange ---
    def get_cells(self):
        probe = Cells()
        probe.conn_ptr = None
        self.L.eo_get_cells(self.h, C.byref(probe))
        n = probe.n
        ptrs = np.zeros(n + 1, dtype=np.int32)
        probe2 = Cells()
        probe2.conn_ptr = ptr(ptrs, _ip)
        self.L.eo_get_cells(self.h, C.byref(probe2))
        nnz = int(ptrs[n])
        out = {k: np.zeros(n, dtype=dt) for k, dt in CELL_DTYPES.items()}
        out["lat"] = np.zeros(3 * n, dtype=np.int32)
        out["conn_ptr"] = np.zeros(n + 1, dtype=np.int32)
        out["conn"] = np.zeros(max(nnz, 1), dtype=np.int32)
        out["gj"] = np.zeros(max(nnz, 1), dtype=np.float64)
        out["memb_adj"] = np.zeros(6 * probe.nmemb, dtype=np.int32)
        s = Cells()
        for k, dt in CELL_DTYPES.items():
            t = {np.uint8: _u8p, np.int32: _ip, np.float64: _dp, np.int64: _i64p}[dt]
            setattr(s, k, ptr(out[k], t))
        s.lat = ptr(out["lat"], _ip)
        s.conn_ptr = ptr(out["conn_ptr"], _ip)
        s.conn = ptr(out["conn"], _ip)
        s.gj = ptr(out["gj"], _dp)
        s.conn_cap = out["conn"].size
        s.memb_adj = ptr(out["memb_adj"], _ip)
        rc = self.L.eo_get_cells(self.h, C.byref(s))
        if rc:
            raise RuntimeError(f"eo_get_cells {rc}")
        out["conn"] = out["conn"][:nnz]
        out["gj"] = out["gj"][:nnz]
        out["n"], out["nmemb"], out["nder"], out["sw"] = s.n, s.nmemb, s.nder, s.sw
        return out

    def get_fields(self):
        sz = int(np.prod(self.shape))
        o = {k: np.zeros(sz, dtype=np.float64) for k in ("atp", "atp_next", "b", "b_next")}
        o["cmap1"] = np.zeros(sz, dtype=np.int32)
        o["cmap2"] = np.zeros(sz, dtype=np.uint8)
        self.L.eo_get_fields(self.h, ptr(o["atp"], _dp), ptr(o["atp_next"], _dp),
                             ptr(o["b"], _dp), ptr(o["b_next"], _dp),
                             ptr(o["cmap1"], _ip), ptr(o["cmap2"], _u8p))
        return {k: v.reshape(self.shape) for k, v in o.items()}

    def set_fields(self, atp=None, atp_next=None, b=None, b_next=None):
        cv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
        a, an, bb, bn = map(cv, (atp, atp_next, b, b_next))
        self._keep = (a, an, bb, bn)
        self.L.eo_set_fields(self.h, ptr(a, _dp), ptr(an, _dp), ptr(bb, _dp), ptr(bn, _dp))

    def set_maps(self, cmap1=None, cmap2=None, lat=None):
        c1 = None if cmap1 is None else np.ascontiguousarray(cmap1, dtype=np.int32).ravel()
        c2 = None if cmap2 is None else np.ascontiguousarray(cmap2, dtype=np.uint8).ravel()
        la = None if lat is None else np.ascontiguousarray(lat, dtype=np.int32).ravel()
        self.L.eo_set_maps(self.h, ptr(c1, _ip), ptr(c2, _u8p), ptr(la, _ip))

    def set_positions(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        self.L.eo_set_positions(self.h, ptr(x, _dp), ptr(y, _dp), ptr(z, _dp))

    def set_state(self, state=None, agek=None):
        s = None if state is None else np.ascontiguousarray(state, dtype=np.uint8)
        a = None if agek is None else np.ascontiguousarray(agek, dtype=np.float64)
        self.L.eo_set_state(self.h, ptr(s, _u8p), ptr(a, _dp))

    def set_lists(self, conn_ptr, conn, gj=None):
        cp = np.ascontiguousarray(conn_ptr, dtype=np.int32)
        cn = np.ascontiguousarray(conn, dtype=np.int32)
        g = None if gj is None else np.ascontiguousarray(gj, dtype=np.float64)
        rc = self.L.eo_set_lists(self.h, ptr(cp, _ip), ptr(cn, _ip), ptr(g, _dp))
        if rc:
            raise RuntimeError(f"eo_set_lists {rc}")

    def p_diff_x(self, a, b):
        return float(self.L.eo_p_diff_x(self.h, a, b))

    def coef(self, law, c1, c2):
        return float(self.L.eo_coef(self.h, law, c1, c2))

    def tri_bend(self, m):
        out = np.zeros(3)
        self.L.eo_tri_bend(self.h, m, ptr(out, _dp))
        return out


LAWS = {"NONE": 0, "OTHER": 1, "DER_DER": 2, "FIXMU": 3, "FIX_MEMB": 4,
        "MU_MEMB": 5, "ALAIRDE": 6, "MEMB_MEMB": 7, "PAIR": 8}
