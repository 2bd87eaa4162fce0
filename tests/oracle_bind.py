"""ctypes binding of the CPU oracle (oracle/libmoseoracle.so) for the tests -- the checker, never the product."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE_DIR = REPO / "oracle"
ORACLE_LIB = ORACLE_DIR / "libmoseoracle.so"

_L = None


def oracle_lib():
    global _L
    if _L is not None:
        return _L
    if not ORACLE_LIB.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
    L = C.CDLL(str(ORACLE_LIB))
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    L.oa_create.restype = vp
    L.oa_create.argtypes = [vp, vp, ip, C.POINTER(C.c_int), dp, dp, dp, dp, dp, dp, dp, ip, vp, ip, vp, ip]
    L.oa_destroy.argtypes = [vp]
    L.oa_ncell.restype = C.c_int64
    L.oa_ncell.argtypes = [vp]
    L.oa_fill_ghosts.argtypes = [vp, dp]
    L.oa_residual.argtypes = [vp, dp, dp, C.c_double]
    L.oa_dt.restype = C.c_double
    L.oa_dt.argtypes = [vp, dp]
    L.oa_newstate.argtypes = [vp, dp, dp, dp, C.c_double, C.c_double]
    L.oa_irs.argtypes = [vp, dp, C.c_double]
    L.oa_step.restype = C.c_int
    L.oa_step.argtypes = [vp, dp, ip, dp]
    L.oa_limiter.restype = C.c_double
    L.oa_limiter.argtypes = [ip, C.c_double, C.c_double]
    L.oa_conv_face.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.c_double, dp]
    L.oa_riemann.argtypes = [vp, ip, dp, dp, dp, dp]
    L.oa_set_chemistry.argtypes = [vp, vp, ip, dp, dp]
    L.oa_chem_failed.restype = C.c_int64
    L.oa_chem_failed.argtypes = [vp]
    L.oa_react_only.argtypes = [vp, dp, C.c_double]
    L.oa_sound_speed.restype = C.c_double
    L.oa_sound_speed.argtypes = [vp, dp, C.c_double]
    L.oc_create.restype = vp
    L.oc_create.argtypes = [vp]
    L.oc_destroy.argtypes = [vp]
    L.oc_rhs.argtypes = [vp, dp, dp]
    L.oc_wdot_pub.argtypes = [vp, dp, C.c_double, dp]
    L.oc_dec.restype = C.c_int
    L.oc_dec.argtypes = [C.c_int, dp, C.POINTER(C.c_int)]
    L.oc_sol.argtypes = [C.c_int, dp, dp, C.POINTER(C.c_int)]
    L.oc_rodas_robertson.restype = C.c_int
    L.oc_rodas_robertson.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int)]
    ip_ = C.POINTER(C.c_int)
    L.oc_decc.restype = C.c_int
    L.oc_decc.argtypes = [C.c_int, dp, dp, ip_]
    L.oc_solc.argtypes = [C.c_int, dp, dp, dp, dp, ip_]
    L.oc_radau5_robertson.restype = C.c_int
    L.oc_radau5_robertson.argtypes = [dp, C.c_double, C.c_double, C.c_double, ip_]
    L.oc_ode_linear.restype = C.c_int
    L.oc_ode_linear.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, ip_]
    L.oc_radau5_tols.argtypes = [C.c_int, dp, dp, dp, dp, dp]
    L.oc_radau5_consts.argtypes = [dp]
    L.oc_advance_m.restype = C.c_int
    L.oc_advance_m.argtypes = [vp, C.c_int, C.c_int64, dp, C.c_double, dp, dp, dp, C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32), C.c_int]
    L.oc_set_reference_pow.argtypes = [C.c_int]
    L.oc_pow_batch.argtypes = [C.c_int, C.c_int64, dp, dp, dp]
    L.oc_advance.restype = C.c_int
    L.oc_advance.argtypes = [vp, C.c_int64, dp, C.c_double, dp, dp, dp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                             C.c_int]
    _L = L
    return L


def _d(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class OracleCase:
    """The oracle instantiated on the same scheme, tables, geometry and records as a product Case."""

    def __init__(self, case, nthreads: int = 8):
        L = oracle_lib()
        self.case = case
        self._t = case.tables.as_struct()
        g = case.geom
        self._keep = (g,)
        self.h = L.oa_create(C.cast(C.pointer(case.scheme), C.c_void_p), C.cast(C.pointer(self._t), C.c_void_p),
                             case.layout.nb, case.layout.dims_arr.ctypes.data_as(C.POINTER(C.c_int)),
                             _d(g.dl), _d(g.m), _d(g.vol), _d(g.yn), _d(g.fm1), _d(g.fm2), _d(g.fm3),
                             len(case.records), case.records.ctypes.data_as(C.c_void_p),
                             len(case.farfield), case.farfield.ctypes.data_as(C.c_void_p), nthreads)
        assert self.h, "oa_create failed"

    def step(self, P: np.ndarray, nsteps: int):
        dts = np.zeros(max(nsteps, 1))
        oracle_lib().oa_step(self.h, _d(P), nsteps, _d(dts))
        return dts[:nsteps]

    def residual(self, P: np.ndarray, dt: float = 0.0) -> np.ndarray:
        R = np.zeros_like(P)
        oracle_lib().oa_residual(self.h, _d(P), _d(R), dt)
        return R

    def newstate(self, P: np.ndarray, PO: np.ndarray, R: np.ndarray, dt: float, rkc: float):
        """one RK stage update of P from the base PO and the residual R (Newstate_Conservative_blk)"""
        oracle_lib().oa_newstate(self.h, _d(P), _d(PO), _d(R), dt, rkc)

    def irs(self, R: np.ndarray, dt: float):
        """the implicit residual smoothing of R in place (IRS_apply): the scaling for the conservative update, the
        residual's ghost layer and the Jacobi sweeps of every computational direction"""
        oracle_lib().oa_irs(self.h, _d(R), dt)

    def dt(self, P: np.ndarray) -> float:
        return oracle_lib().oa_dt(self.h, _d(P))

    def fill_ghosts(self, P: np.ndarray):
        oracle_lib().oa_fill_ghosts(self.h, _d(P))

    def riemann(self, rs: int, qL, qR, n):
        F = np.zeros(5)
        oracle_lib().oa_riemann(self.h, rs, _d(np.ascontiguousarray(qL, dtype=float)),
                                _d(np.ascontiguousarray(qR, dtype=float)), _d(np.ascontiguousarray(n, dtype=float)),
                                _d(F))
        return F

    def conv_face(self, pm1, p0, pp1, pp2, dlc, nrm, area):
        res = np.zeros(len(p0))
        a = [np.ascontiguousarray(x, dtype=float) for x in (pm1, p0, pp1, pp2, dlc, nrm)]
        oracle_lib().oa_conv_face(self.h, *[_d(x) for x in a], float(area), _d(res))
        return res

    def set_chemistry(self, ochem: "OracleChem | None", method: int = 1, rtol: float = 1e-8, atol: float = 1e-12):
        """Strang-split steps with the chemistry oracle's mechanism (None: plain steps), as mose_case_chem_attach."""
        nz = self.case.scheme.nsc + 1
        self._chem = ochem
        self._tol = (np.full(nz, rtol), np.full(nz, atol))
        oracle_lib().oa_set_chemistry(self.h, None if ochem is None else ochem.h, method, _d(self._tol[0]),
                                      _d(self._tol[1]))

    def chem_failed(self) -> int:
        return oracle_lib().oa_chem_failed(self.h)

    def react(self, P: np.ndarray, dt: float):
        """Chemistry_Newstate_blk alone on the interior of P (the chemistry of set_chemistry)."""
        oracle_lib().oa_react_only(self.h, _d(P), dt)

    def sound_speed(self, rhoi, p):
        return oracle_lib().oa_sound_speed(self.h, _d(np.ascontiguousarray(rhoi, dtype=float)), float(p))

    def __del__(self):
        try:
            oracle_lib().oa_destroy(self.h)
        except Exception:
            pass


class OracleChem:
    """The chemistry oracle (oracle/mose_chem_oracle.c) on a mechanism of mosegpu.chem."""

    def __init__(self, mech, nthreads: int = 8):
        self.mech = mech
        self._m = mech.as_struct()
        self.h = oracle_lib().oc_create(C.cast(C.pointer(self._m), C.c_void_p))
        assert self.h, "oc_create failed"
        self.nthreads = nthreads

    def rhs(self, Z):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        F = np.zeros_like(Z)
        oracle_lib().oc_rhs(self.h, _d(Z), _d(F))
        return F

    def wdot(self, roi, T):
        roi = np.ascontiguousarray(roi, dtype=np.float64)
        om = np.zeros_like(roi)
        oracle_lib().oc_wdot_pub(self.h, _d(roi), T, _d(om))
        return om

    def advance(self, state, dt, rtol=1e-8, atol=1e-12, method=1):
        """method 1: RODAS, 2: RADAU5 (the MOSE_ODE_* ids)"""
        st = np.ascontiguousarray(state, dtype=np.float64).copy()
        n = st.shape[1]
        nz = self.mech.ns + 1
        rt = np.full(nz, rtol); at = np.full(nz, atol)
        hr = np.zeros(n)
        idid = np.zeros(n, dtype=np.int32)
        nstep = np.zeros(n, dtype=np.int32)
        oracle_lib().oc_advance_m(self.h, method, n, _d(st), dt, _d(rt), _d(at), _d(hr),
                                  idid.ctypes.data_as(C.POINTER(C.c_int32)),
                                  nstep.ctypes.data_as(C.POINTER(C.c_int32)), self.nthreads)
        return st, hr, idid, nstep


REF_LIB = ORACLE_DIR / "_ref" / "libhairer_ref.so"
_R = None


def ref_lib():
    """The reference's own Hairer RODAS / RADAU5, compiled from the reference sources by `make -C oracle ref` (absent
    where the reference tree is not: the tests that need it skip)."""
    global _R
    if _R is not None:
        return _R
    if not REF_LIB.exists():
        return None
    R = C.CDLL(str(REF_LIB))
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    R.ref_robertson.restype = C.c_int
    R.ref_robertson.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, ip]
    R.ref_advance_cell.restype = C.c_int
    R.ref_advance_cell.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, C.c_double, dp, dp, dp, ip]
    _R = R
    return R
