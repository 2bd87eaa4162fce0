"""ctypes binding of the MI355X C ABI (include/mose_gpu.h).

The library is the in-tree ``mose-gpu_amd/lib/libmosegpu.so``; every entry point raises ``MoseError`` on a
negative status, with the context's message.  Nothing here falls back to a CPU path: without the library or a
GPU the calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent          # mose-gpu_amd/
LIB_PATH = PKG_DIR / "lib" / "libmosegpu.so"

GC = 2
NTAB = 10001

MOSE_OK = 0
ERRORS = {
    -1: "MOSE_ERR_ARG", -2: "MOSE_ERR_HIP", -3: "MOSE_ERR_NOMEM", -4: "MOSE_ERR_MESH",
    -5: "MOSE_ERR_STATE", -6: "MOSE_ERR_THERMO", -7: "MOSE_ERR_NODEV", -8: "MOSE_ERR_ODE",
}

BC_NONE, BC_CONNECTION, BC_SYMMETRY, BC_FARFIELD, BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL, BC_EXTRAPOLATION = 0, 1, 3, 4, 5, 6, 11
LIM = {"minmod": 1, "vanleer": 2, "vanalbada": 3, "mc": 4, "superbee": 5, "first-order": 6}
RS = {"HLLC": 1, "HLLE": 2, "SLAU": 3, "AUSM+": 4, "AUSMp": 4, "LMRoe": 12, "Roe": 12}
# integration-variables of the state update (Lib_Newstate.f90:43-60): conservative or primitive
INTEGRATION = {"cons": 0, "prim": 1}
INT_CONS, INT_PRIM = 0, 1

# one boundary cell-face (mose_bc_record)
RECORD_DTYPE = np.dtype(
    [("b", "<i4"), ("i", "<i4"), ("j", "<i4"), ("k", "<i4"), ("f", "<i4"), ("typ", "<i4"),
     ("bs", "<i4"), ("is_", "<i4"), ("js", "<i4"), ("ks", "<i4"), ("fs", "<i4"),
     ("d11", "<i4"), ("d12", "<i4"), ("d21", "<i4"), ("d22", "<i4"), ("periodic", "<i4"),
     ("ff", "<i4"), ("pad", "<i4"), ("tw", "<f8"), ("tvec", "<f8", (3,))],
    align=True,
)
assert RECORD_DTYPE.itemsize == 104

MAX_SPECIES = 16
# the data of a far-field / inflow-outflow face (mose_farfield)
FARFIELD_DTYPE = np.dtype(
    [("mach", "<f8"), ("t0", "<f8"), ("p0", "<f8"), ("pamb", "<f8"), ("alpha", "<f8"), ("beta", "<f8"),
     ("rel_fac", "<f8"), ("reserved", "<f8"), ("ci", "<f8", (MAX_SPECIES,)), ("rans", "<f8", (2,))],
    align=True,
)
assert FARFIELD_DTYPE.itemsize == 208


class MoseError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Scheme(C.Structure):
    _fields_ = [
        ("nsc", C.c_int32), ("nrans", C.c_int32), ("limiter", C.c_int32), ("riemann", C.c_int32),
        ("viscous", C.c_int32), ("n_rk", C.c_int32), ("ndir", C.c_int32), ("point_implicit", C.c_int32),
        ("thermo_guard", C.c_int32), ("integration", C.c_int32), ("irs", C.c_int32), ("irs_njacobi", C.c_int32),
        ("reserved", C.c_int32 * 4),
        ("mco", C.c_double), ("cfl", C.c_double), ("vnn", C.c_double), ("prandtl", C.c_double),
        ("sc", C.c_double), ("sct", C.c_double), ("prt", C.c_double), ("irs_beta", C.c_double),
        ("reserved_d", C.c_double * 7),
    ]


class Tables(C.Structure):
    _fields_ = [("nsc", C.c_int32), ("pad", C.c_int32), ("wm", C.POINTER(C.c_double)),
                ("cp", C.POINTER(C.c_double)), ("h", C.POINTER(C.c_double)),
                ("mu", C.POINTER(C.c_double)), ("k", C.POINTER(C.c_double))]


PATH_FUSED, PATH_PLAIN, PATH_FAST = 0, 1, 2


class Tuning(C.Structure):
    _fields_ = [("path", C.c_int32), ("ti", C.c_int32), ("tj", C.c_int32), ("tk", C.c_int32),
                ("timing", C.c_int32), ("no_graph", C.c_int32), ("reserved", C.c_int32 * 2)]


class Report(C.Structure):
    _fields_ = [("dt", C.c_double), ("time", C.c_double), ("steps", C.c_int64), ("bad_cells", C.c_int32),
                ("first_bad_block", C.c_int32), ("first_bad_i", C.c_int32), ("first_bad_j", C.c_int32),
                ("first_bad_k", C.c_int32), ("status", C.c_int32)]


class MechanismStruct(C.Structure):
    _fields_ = [("ns", C.c_int32), ("n_arrh", C.c_int32), ("n_troe", C.c_int32), ("n_lind", C.c_int32),
                ("tmin", C.c_int32), ("tmax", C.c_int32),
                ("wm", C.POINTER(C.c_double)), ("cp", C.POINTER(C.c_double)), ("h", C.POINTER(C.c_double)),
                ("ni1", C.POINTER(C.c_double)), ("ni2", C.POINTER(C.c_double)), ("eps", C.POINTER(C.c_double)),
                ("kf", C.POINTER(C.c_double)), ("kb", C.POINTER(C.c_double)), ("troe", C.POINTER(C.c_double)),
                ("lind", C.POINTER(C.c_double))]


class OdeOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("reserved", C.c_int32 * 3), ("rtol", C.c_double * 17),
                ("atol", C.c_double * 17)]


class ChemReport(C.Structure):
    _fields_ = [("cells", C.c_int64), ("failed", C.c_int64), ("steps_total", C.c_int64),
                ("rejected_total", C.c_int64), ("rhs_total", C.c_int64), ("steps_max", C.c_int32),
                ("first_failed_idid", C.c_int32), ("first_failed_cell", C.c_int64), ("ms", C.c_double)]


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


_lib_handle = None


def lib() -> C.CDLL:
    """The loaded C ABI.  Raises if the in-tree library is not built."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not LIB_PATH.exists():
        raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(str(LIB_PATH))
    vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    L.mose_last_error.restype = C.c_char_p
    L.mose_last_error.argtypes = [vp]
    L.mose_device_count.argtypes = [C.POINTER(C.c_int)]
    L.mose_ctx_create.argtypes = [ip, C.POINTER(Tables), C.POINTER(vp)]
    L.mose_ctx_destroy.argtypes = [vp]
    L.mose_metrics_build.argtypes = [ip, C.POINTER(C.c_int), dp, ip, C.c_void_p, ip, dp, dp, dp, dp, dp, dp, dp]
    L.mose_records_tangent.argtypes = [ip, C.POINTER(C.c_int), dp, ip, C.c_void_p]
    L.mose_wall_distance.argtypes = [ip, C.POINTER(C.c_int), dp, ip, C.c_void_p, dp]
    L.mose_case_create.argtypes = [vp, C.POINTER(Scheme), ip, C.POINTER(C.c_int), dp, dp, dp, dp, dp, dp, dp, ip,
                                   C.c_void_p, ip, C.c_void_p, C.POINTER(C.c_int)]
    L.mose_case_destroy.argtypes = [vp, ip]
    L.mose_case_bytes.argtypes = [C.POINTER(Scheme), ip, C.POINTER(C.c_int), ip, C.POINTER(C.c_int64)]
    L.mose_state_upload.argtypes = [vp, ip, dp]
    L.mose_state_download.argtypes = [vp, ip, dp]
    L.mose_case_step.argtypes = [vp, ip, ip, C.POINTER(Report)]
    L.mose_cases_step.argtypes = [vp, C.POINTER(C.c_int), ip, ip, C.POINTER(Report)]
    L.mose_case_launch.argtypes = [vp, ip, ip]
    L.mose_case_poll.argtypes = [vp, ip, C.POINTER(C.c_int), C.POINTER(Report)]
    L.mose_case_residual.argtypes = [vp, ip, dp]
    L.mose_case_dt.argtypes = [vp, ip, dp]
    L.mose_case_residual_norm.argtypes = [vp, ip, dp]
    L.mose_case_state_hash.argtypes = [vp, ip, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.mose_case_timing.argtypes = [vp, ip, dp, dp, C.POINTER(C.c_int64)]
    L.mose_case_tune.argtypes = [vp, ip, C.POINTER(Tuning), C.POINTER(Tuning)]
    L.mose_bandwidth_probe.argtypes = [vp, C.c_int64, ip, dp]
    L.mose_ctx_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mose_chem_create.argtypes = [vp, C.POINTER(MechanismStruct), C.c_int64, C.POINTER(C.c_int)]
    L.mose_chem_destroy.argtypes = [vp, ip]
    L.mose_chem_upload.argtypes = [vp, ip, C.c_int64, dp]
    L.mose_chem_download.argtypes = [vp, ip, dp, dp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.mose_chem_advance.argtypes = [vp, ip, C.c_double, C.POINTER(OdeOpts), C.POINTER(ChemReport)]
    L.mose_pow_batch.argtypes = [vp, C.c_int64, dp, dp, dp]
    L.mose_chem_advance_cells.argtypes = [vp, ip, C.c_int64, C.c_int64, vp, vp, vp, vp, C.c_double,
                                          C.POINTER(OdeOpts), C.POINTER(ChemReport)]
    L.mose_case_chem_attach.argtypes = [vp, ip, ip, C.POINTER(OdeOpts)]
    L.mose_case_chem_advance.argtypes = [vp, ip, C.c_double, C.POINTER(ChemReport)]
    L.mose_case_chem_report.argtypes = [vp, ip, C.POINTER(ChemReport)]
    for name in ("mose_ctx_create", "mose_ctx_destroy", "mose_metrics_build", "mose_records_tangent", "mose_case_create",
                 "mose_case_destroy", "mose_case_bytes", "mose_state_upload", "mose_state_download",
                 "mose_case_step", "mose_cases_step", "mose_case_launch", "mose_case_poll", "mose_case_residual",
                 "mose_case_dt",
                 "mose_case_residual_norm", "mose_case_state_hash", "mose_case_timing", "mose_device_count",
                 "mose_case_tune", "mose_bandwidth_probe", "mose_ctx_info", "mose_chem_create", "mose_chem_destroy",
                 "mose_chem_upload", "mose_chem_download", "mose_chem_advance", "mose_pow_batch",
                 "mose_chem_advance_cells", "mose_case_chem_attach", "mose_case_chem_advance",
                 "mose_case_chem_report", "mose_wall_distance"):
        getattr(L, name).restype = C.c_int
    _lib_handle = L
    return L


def device_count() -> int:
    n = C.c_int(0)
    lib().mose_device_count(C.byref(n))
    return n.value


# ------------------------------------------------------------------------------------------------------------
# geometry
# ------------------------------------------------------------------------------------------------------------
def state_hash(layout: "Layout", P: np.ndarray) -> tuple[int, int]:
    """field_state_hash (MOSE_Field_m.f90:254-284) of a host state: the wrapping 64-bit sums of the low and of the
    high 32 bits of the IEEE patterns of every interior value."""
    lo = hi = 0
    for b in range(layout.nb):
        x = np.ascontiguousarray(layout.interior(P, b)).view(np.uint64)
        lo += int(np.sum(x & np.uint64(0xFFFFFFFF), dtype=np.uint64))
        hi += int(np.sum(x >> np.uint64(32), dtype=np.uint64))
    return lo % (1 << 64), hi % (1 << 64)


class Layout:
    """Arena sizes of a list of blocks (rows of the padded cells, face counts)."""

    def __init__(self, dims):
        self.dims = [tuple(int(x) for x in d) for d in dims]
        self.nb = len(self.dims)
        self.boff, self.f1off, self.f2off, self.f3off = [], [], [], []
        NC = NF1 = NF2 = NF3 = 0
        for n1, n2, n3 in self.dims:
            self.boff.append(NC); self.f1off.append(NF1); self.f2off.append(NF2); self.f3off.append(NF3)
            NC += (n1 + 2 * GC) * (n2 + 2 * GC) * (n3 + 2 * GC)
            NF1 += (n1 + 1) * n2 * n3
            NF2 += n1 * (n2 + 1) * n3
            NF3 += n1 * n2 * (n3 + 1)
        self.NC, self.NF1, self.NF2, self.NF3 = NC, NF1, NF2, NF3
        self.dims_arr = np.ascontiguousarray(np.array(self.dims, dtype=np.int32).reshape(-1))

    def block_view(self, arr: np.ndarray, b: int, nvar: int | None = None):
        """View of block b of an arena array [nvar, NC] as [nvar, n3+4, n2+4, n1+4] (k, j, i)."""
        n1, n2, n3 = self.dims[b]
        shp = (n3 + 2 * GC, n2 + 2 * GC, n1 + 2 * GC)
        npad = shp[0] * shp[1] * shp[2]
        if arr.ndim == 1:
            return arr[self.boff[b]:self.boff[b] + npad].reshape(shp)
        return arr[:, self.boff[b]:self.boff[b] + npad].reshape((arr.shape[0],) + shp)

    def interior(self, arr: np.ndarray, b: int):
        v = self.block_view(arr, b)
        return v[..., GC:-GC, GC:-GC, GC:-GC]


class Geometry:
    def __init__(self, layout: Layout, dl, m, vol, yn, fm1, fm2, fm3):
        self.layout = layout
        self.dl, self.m, self.vol, self.yn, self.fm1, self.fm2, self.fm3 = dl, m, vol, yn, fm1, fm2, fm3


def metrics_build(layout: Layout, nodes: list[np.ndarray], records: np.ndarray, ndir: int) -> Geometry:
    """Setup_Metrics (Lib_Metrics.f90:15-136) through mose_metrics_build.  nodes[b]: (n3+1, n2+1, n1+1, 3)."""
    flat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(n, dtype=np.float64).reshape(-1) for n in nodes]))
    NC = layout.NC
    dl = np.zeros(3 * NC); m = np.zeros(9 * NC); vol = np.zeros(NC); yn = np.zeros(NC)
    fm1 = np.zeros(4 * layout.NF1); fm2 = np.zeros(4 * layout.NF2); fm3 = np.zeros(4 * layout.NF3)
    rec = np.ascontiguousarray(records, dtype=RECORD_DTYPE)
    rc = lib().mose_metrics_build(layout.nb, iptr(layout.dims_arr), dptr(flat), len(rec),
                                  rec.ctypes.data_as(C.c_void_p), ndir, dptr(dl), dptr(m), dptr(vol), dptr(yn),
                                  dptr(fm1), dptr(fm2), dptr(fm3))
    if rc != MOSE_OK:
        raise MoseError(rc, "mose_metrics_build failed (degenerate cell or bad record)")
    return Geometry(layout, dl, m, vol, yn, fm1, fm2, fm3)


def wall_distance(layout: Layout, nodes: list[np.ndarray], records: np.ndarray, yn: np.ndarray) -> None:
    """Compute_Yn + Yn_Connection (Lib_Metrics.f90:513-763) through mose_wall_distance, into yn [NC] in place."""
    flat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(n, dtype=np.float64).reshape(-1) for n in nodes]))
    rec = np.ascontiguousarray(records, dtype=RECORD_DTYPE)
    assert yn.dtype == np.float64 and yn.flags.c_contiguous and yn.size == layout.NC
    rc = lib().mose_wall_distance(layout.nb, iptr(layout.dims_arr), dptr(flat), len(rec),
                                  rec.ctypes.data_as(C.c_void_p), dptr(yn))
    if rc != MOSE_OK:
        raise MoseError(rc, "mose_wall_distance failed")


def records_tangent(layout: Layout, nodes: list[np.ndarray], records: np.ndarray) -> None:
    """The unit face tangent of the far-field records, in place (Setup_BC_Device, Mod_BC.f90:228-240)."""
    flat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(n, dtype=np.float64).reshape(-1) for n in nodes]))
    assert records.dtype == RECORD_DTYPE and records.flags.c_contiguous
    rc = lib().mose_records_tangent(layout.nb, iptr(layout.dims_arr), dptr(flat), len(records),
                                    records.ctypes.data_as(C.c_void_p))
    if rc != MOSE_OK:
        raise MoseError(rc, "mose_records_tangent failed")


# ------------------------------------------------------------------------------------------------------------
# context and cases
# ------------------------------------------------------------------------------------------------------------
class ThermoTables:
    """FLINT per-Kelvin tables on T = 0..10000 K (Load_ThermoTransport.f90:7-189): cp, h, mu, k [nsc, 10001]."""

    def __init__(self, wm, cp, h, mu=None, k=None, names=None):
        self.wm = np.ascontiguousarray(np.atleast_1d(np.asarray(wm, dtype=np.float64)))
        self.nsc = self.wm.size
        self.cp = np.ascontiguousarray(np.asarray(cp, dtype=np.float64).reshape(self.nsc, NTAB))
        self.h = np.ascontiguousarray(np.asarray(h, dtype=np.float64).reshape(self.nsc, NTAB))
        self.mu = None if mu is None else np.ascontiguousarray(np.asarray(mu, dtype=np.float64).reshape(self.nsc, NTAB))
        self.k = None if k is None else np.ascontiguousarray(np.asarray(k, dtype=np.float64).reshape(self.nsc, NTAB))
        self.names = names or [f"s{q}" for q in range(self.nsc)]

    def as_struct(self) -> Tables:
        t = Tables()
        t.nsc = self.nsc
        t.wm = dptr(self.wm); t.cp = dptr(self.cp); t.h = dptr(self.h)
        t.mu = dptr(self.mu); t.k = dptr(self.k)
        return t


class Context:
    def __init__(self, tables: ThermoTables, device: int = 0):
        self.tables = tables
        self._t = tables.as_struct()
        h = C.c_void_p()
        rc = lib().mose_ctx_create(device, C.byref(self._t), C.byref(h))
        if rc != MOSE_OK:
            raise MoseError(rc, "mose_ctx_create failed")
        self.h = h

    def check(self, rc, what):
        if rc != MOSE_OK:
            raise MoseError(rc, f"{what}: {lib().mose_last_error(self.h).decode()}")
        return rc

    def info(self) -> dict:
        t, f, n, l = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
        self.check(lib().mose_ctx_info(self.h, C.byref(t), C.byref(f), C.byref(n), C.byref(l)), "mose_ctx_info")
        return {"hbm_total": t.value, "hbm_free": f.value, "cus": n.value, "lds_per_block": l.value}

    def bandwidth_probe(self, nbytes: int = 1 << 30, reps: int = 20) -> float:
        """GB/s (read + write) of the library's stream-copy kernel on `nbytes` buffers."""
        v = C.c_double(0.0)
        self.check(lib().mose_bandwidth_probe(self.h, int(nbytes), int(reps), C.byref(v)), "mose_bandwidth_probe")
        return v.value

    def pow_batch(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """x**y elementwise on the device with the chemistry kernels' pow (the host libm's bits)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        if x.shape != y.shape:
            raise ValueError("pow_batch: x and y differ in shape")
        out = np.empty_like(x)
        self.check(lib().mose_pow_batch(self.h, x.size, dptr(x), dptr(y), dptr(out)), "mose_pow_batch")
        return out

    def close(self):
        if self.h:
            lib().mose_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCase:
    """A case resident on the device (mose_case_create + upload/step/download)."""

    def __init__(self, ctx: Context, case):
        self.ctx = ctx
        self.case = case
        g = case.geom
        cid = C.c_int(-1)
        ctx.check(lib().mose_case_create(ctx.h, C.byref(case.scheme), case.layout.nb, iptr(case.layout.dims_arr),
                                         dptr(g.dl), dptr(g.m), dptr(g.vol), dptr(g.yn), dptr(g.fm1), dptr(g.fm2),
                                         dptr(g.fm3), len(case.records), case.records.ctypes.data_as(C.c_void_p),
                                         len(case.farfield), case.farfield.ctypes.data_as(C.c_void_p), C.byref(cid)),
                  "mose_case_create")
        self.id = cid.value

    def upload(self, P: np.ndarray):
        P = np.ascontiguousarray(P, dtype=np.float64)
        self.ctx.check(lib().mose_state_upload(self.ctx.h, self.id, dptr(P)), "mose_state_upload")

    def download(self) -> np.ndarray:
        P = np.zeros((self.case.nprim, self.case.layout.NC))
        self.ctx.check(lib().mose_state_download(self.ctx.h, self.id, dptr(P)), "mose_state_download")
        return P

    def step(self, nsteps: int) -> Report:
        rep = Report()
        self.ctx.check(lib().mose_case_step(self.ctx.h, self.id, nsteps, C.byref(rep)), "mose_case_step")
        return rep

    def run_until(self, t_end: float, max_steps: int = 10 ** 7) -> Report:
        """Steps while the case's time is below t_end: the reference's time-threshold loop (`do while time <
        timemax`, without the optional dt clip of Clip_dt_To_Timemax, Mod_dt.f90:283-296), so the last step may end
        past t_end.  One step per call: the time of each step is the device's dt of that step."""
        rep = self.step(0)
        n = 0
        while rep.time < t_end and n < max_steps:
            rep = self.step(1)
            n += 1
        return rep

    def launch(self, nsteps: int):
        """Queue nsteps steps and return at once (mose_case_launch); poll() until it reports them done."""
        self.ctx.check(lib().mose_case_launch(self.ctx.h, self.id, nsteps), "mose_case_launch")

    def poll(self) -> Report | None:
        """The report of the launched steps once they have finished, None while they run (mose_case_poll)."""
        done = C.c_int(0)
        rep = Report()
        self.ctx.check(lib().mose_case_poll(self.ctx.h, self.id, C.byref(done), C.byref(rep)), "mose_case_poll")
        return rep if done.value else None

    def residual(self) -> np.ndarray:
        R = np.zeros((self.case.nprim, self.case.layout.NC))
        self.ctx.check(lib().mose_case_residual(self.ctx.h, self.id, dptr(R)), "mose_case_residual")
        return R

    def dt(self) -> float:
        v = C.c_double(0.0)
        self.ctx.check(lib().mose_case_dt(self.ctx.h, self.id, C.byref(v)), "mose_case_dt")
        return v.value

    def residual_norm(self) -> np.ndarray:
        """sqrt of the summed squared changes of the last step: species sum, u, v, w, p, RANS (nprim - nsc + 1)."""
        out = np.zeros(self.case.nprim - self.case.scheme.nsc + 1)
        self.ctx.check(lib().mose_case_residual_norm(self.ctx.h, self.id, dptr(out)), "mose_case_residual_norm")
        return out

    def state_hash(self) -> tuple[int, int]:
        """field_state_hash: (sum of the low, sum of the high 32-bit halves) of every interior value."""
        lo, hi = C.c_uint64(0), C.c_uint64(0)
        self.ctx.check(lib().mose_case_state_hash(self.ctx.h, self.id, C.byref(lo), C.byref(hi)), "mose_case_state_hash")
        return lo.value, hi.value

    def tune(self, path: int = PATH_FUSED, tile=(0, 0, 0), timing: bool = False, graph: bool = True):
        """Execution path and tile (0,0,0 = autotune); graph=False launches every fused step eagerly.  Returns the
        tile chosen."""
        t, out = Tuning(), Tuning()
        t.path = path; t.ti, t.tj, t.tk = tile; t.timing = int(timing); t.no_graph = int(not graph)
        self.ctx.check(lib().mose_case_tune(self.ctx.h, self.id, C.byref(t), C.byref(out)), "mose_case_tune")
        return (out.ti, out.tj, out.tk)

    def timing(self):
        """(average ms per stage kernel, total ms, launches) of the last timed step call."""
        a, t, n = C.c_double(), C.c_double(), C.c_int64()
        self.ctx.check(lib().mose_case_timing(self.ctx.h, self.id, C.byref(a), C.byref(t), C.byref(n)), "timing")
        return a.value, t.value, n.value

    def chem_attach(self, chem: "DeviceChem | None", opts=None):
        """mose_case_chem_attach: Strang-split steps with chem's mechanism (None: plain transport steps)."""
        rc = lib().mose_case_chem_attach(self.ctx.h, self.id, -1 if chem is None else chem.id,
                                         None if opts is None else C.byref(opts))
        self.ctx.check(rc, "mose_case_chem_attach")

    def chem_advance(self, dt: float) -> "ChemReport":
        """mose_case_chem_advance: Chemistry_Newstate_blk on the resident state over dt."""
        rep = ChemReport()
        self.ctx.check(lib().mose_case_chem_advance(self.ctx.h, self.id, dt, C.byref(rep)), "mose_case_chem_advance")
        return rep

    def chem_report(self) -> "ChemReport":
        rep = ChemReport()
        self.ctx.check(lib().mose_case_chem_report(self.ctx.h, self.id, C.byref(rep)), "mose_case_chem_report")
        return rep

    def destroy(self):
        if self.id >= 0:
            lib().mose_case_destroy(self.ctx.h, self.id)
            self.id = -1


class DeviceChem:
    """A batch of independent chemistry cells resident on the device (mose_chem_*)."""

    def __init__(self, ctx: "Context", mech, ncell_max: int):
        self.ctx = ctx
        self.mech = mech
        self._m = mech.as_struct()
        cid = C.c_int(-1)
        ctx.check(lib().mose_chem_create(ctx.h, C.byref(self._m), int(ncell_max), C.byref(cid)), "mose_chem_create")
        self.id = cid.value
        self.n = 0

    def upload(self, state: np.ndarray):
        st = np.ascontiguousarray(state, dtype=np.float64)
        self.ctx.check(lib().mose_chem_upload(self.ctx.h, self.id, st.shape[1], dptr(st)), "mose_chem_upload")
        self.n = st.shape[1]

    def advance(self, dt: float, opts) -> ChemReport:
        """Raises MoseError(MOSE_ERR_ODE) when cells end with idid < 0; the report of the call, filled either way,
        stays in last_report and the cells' states and idid can still be downloaded."""
        rep = ChemReport()
        self.last_report = rep
        self.ctx.check(lib().mose_chem_advance(self.ctx.h, self.id, dt, C.byref(opts), C.byref(rep)),
                       "mose_chem_advance")
        return rep

    def advance_cells(self, ncell: int, stride: int, state: int, hr: int, idid: int, nstep: int, dt: float,
                      opts) -> ChemReport:
        """mose_chem_advance_cells: the same integration on cells the caller holds in device memory (the integer
        device addresses of state [ns + 1][stride], hr, idid and nstep), with this object's mechanism; the call
        allocates nothing."""
        rep = ChemReport()
        self.last_report = rep
        self.ctx.check(lib().mose_chem_advance_cells(self.ctx.h, self.id, int(ncell), int(stride), state, hr, idid,
                                                     nstep, dt, C.byref(opts), C.byref(rep)),
                       "mose_chem_advance_cells")
        return rep

    def download(self):
        st = np.zeros((self.mech.ns + 1, self.n))
        hr = np.zeros(self.n)
        idid = np.zeros(self.n, dtype=np.int32)
        nstep = np.zeros(self.n, dtype=np.int32)
        self.ctx.check(lib().mose_chem_download(self.ctx.h, self.id, dptr(st), dptr(hr),
                                                idid.ctypes.data_as(C.POINTER(C.c_int32)),
                                                nstep.ctypes.data_as(C.POINTER(C.c_int32))), "mose_chem_download")
        return st, hr, idid, nstep

    def destroy(self):
        if self.id >= 0:
            lib().mose_chem_destroy(self.ctx.h, self.id)
            self.id = -1
