"""epi_calc_ext_stim_noalloc (calc_ext_stim over caller-held device grids,
allocating nothing, ext_stim.cpp:18-69) against epi_calc_ext_stim on the same
tissues: bit-identical B after the sweep, on two calls in a row, the second
with fewer and smaller grids (MI355X only).

Device memory for the caller's grids comes from the HIP runtime the library
itself loaded (the test is the caller: it may allocate; the entry point does
not)."""
import ctypes as C

import numpy as np
import pytest

from epi_amd import configs

pytestmark = pytest.mark.gpu


class View(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("xp", C.c_int32),
                ("pl", C.c_int64), ("lz", C.c_double), ("zzmax", C.c_double),
                ("b_in", C.c_void_p), ("b_out", C.c_void_p), ("cmap1", C.c_void_p), ("cmap2", C.c_void_p),
                ("state", C.c_void_p), ("agek", C.c_void_p)]


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libepi_amd.so uses"""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        self.bufs.append(p)
        return p.value

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 2) == 0   # device to host
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def to_layout(a, nx, ny, nz, xp, pl, fill=0):
    """[x][y][z] array of (nx+1)(ny+1)(nz+1) points into the entry point's
    layout: point (j, k, l) at l pl + k xp + j"""
    a = a.reshape(nx + 1, ny + 1, nz + 1)
    out = np.full(pl * (nz + 1), fill, dtype=a.dtype)
    for l in range(nz + 1):
        plane = out[l * pl:l * pl + xp * (ny + 1)].reshape(ny + 1, xp)
        plane[:, :nx + 1] = a[:, :, l].T
    return out


def from_layout(v, nx, ny, nz, xp, pl):
    out = np.empty((nx + 1, ny + 1, nz + 1), dtype=v.dtype)
    for l in range(nz + 1):
        out[:, :, l] = v[l * pl:l * pl + xp * (ny + 1)].reshape(ny + 1, xp)[:, :nx + 1].T
    return out.ravel()


def specs():
    return [configs.config_tiny(7, n_free=160),
            configs.config_tiny(8, n_free=260, side=20, grid=20, nz=28),
            configs.config_tiny(10, n_free=300, side=24, grid=18, nz=30, dead_top=True, n_layers=3)]


def run_case(hip, lib, spec_list, steps, pad):
    """advance the tissues in a context, then sweep B once through
    epi_calc_ext_stim and once through the no-alloc entry point on copies of
    the same inputs; returns (context result, no-alloc result) per tissue"""
    from epi_amd import native
    ts = native.TissueSet([s["geometry"] for s in spec_list])
    for t, s in enumerate(spec_list):
        ts.load(t, s)
        ts.set_rng(t, native.mt19937_64_uniform(40 + t, 200000))
    ts.step([native.Driver(0, 0, 0, 0, 0) for _ in spec_list], steps)
    zz = ts.calc_zzmax()
    ts.setup_map_lat()
    before = [ts.get_fields(t) for t in range(len(spec_list))]
    cells = [ts.get_cells(t) for t in range(len(spec_list))]
    views = (View * len(spec_list))()
    shapes = []
    for t, s in enumerate(spec_list):
        g = s["geometry"]
        nx, ny, nz = g.nx, g.ny, g.nz
        xp = nx + 1 + pad
        pl = xp * (ny + 1) + 3 * pad
        f = before[t]
        v = views[t]
        v.nx, v.ny, v.nz, v.xp, v.pl, v.lz, v.zzmax = nx, ny, nz, xp, pl, g.lz, float(zz[t])
        v.b_in = hip.put(to_layout(f["b"], nx, ny, nz, xp, pl))
        v.b_out = hip.put(to_layout(f["b_next"], nx, ny, nz, xp, pl))
        v.cmap1 = hip.put(to_layout(f["cmap1"].astype(np.int32), nx, ny, nz, xp, pl, fill=-1))
        v.cmap2 = hip.put(to_layout(f["cmap2"].astype(np.uint8), nx, ny, nz, xp, pl))
        v.state = hip.put(cells[t]["state"].astype(np.uint8))
        v.agek = hip.put(cells[t]["agek"].astype(np.float64))
        shapes.append((nx, ny, nz, xp, pl, v.b_out))
    d_views = hip.put(np.frombuffer(bytes(views), dtype=np.uint8))
    d_err = hip.put(np.full(len(spec_list), -1, dtype=np.int32))
    print(f"no-alloc sweep of {len(spec_list)} grids (pad {pad})", flush=True)
    rc = lib.epi_calc_ext_stim_noalloc(C.c_void_p(d_views), len(spec_list), C.c_void_p(d_err), None)
    assert rc == 0, rc
    hip.sync()
    err = hip.get(d_err, len(spec_list), np.int32)
    got = [from_layout(hip.get(b_out, pl * (nz + 1), np.float64), nx, ny, nz, xp, pl)
           for nx, ny, nz, xp, pl, b_out in shapes]
    ts.calc_ext_stim(zz)
    ref = [ts.get_fields(t)["b"].ravel() for t in range(len(spec_list))]
    ts.close()
    return err, got, ref, before


def test_noalloc_matches_calc_ext_stim():
    from epi_amd import native
    lib = native.load_library()
    lib.epi_calc_ext_stim_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    lib.epi_noalloc_scratch_bytes.restype = C.c_int64
    assert 0 < lib.epi_noalloc_scratch_bytes() <= 4 << 20
    hip = Hip()
    try:
        # limits are refused before anything runs
        assert lib.epi_calc_ext_stim_noalloc(C.c_void_p(8), 0, C.c_void_p(8), None) == 7
        assert lib.epi_calc_ext_stim_noalloc(C.c_void_p(8), 1025, C.c_void_p(8), None) == 7
        for case, (spec_list, steps, pad) in enumerate([(specs(), 6, 5), (specs()[:2][::-1], 3, 0)]):
            err, got, ref, before = run_case(hip, lib, spec_list, steps, pad)
            assert list(err) == [0] * len(spec_list), (case, err)
            for t in range(len(spec_list)):
                assert not np.array_equal(before[t]["b"].ravel(), ref[t]), (case, t)   # the sweep changed B
                assert np.array_equal(got[t], ref[t]), (case, t, float(np.max(np.abs(got[t] - ref[t]))))
            hip.free()
    finally:
        hip.free()
