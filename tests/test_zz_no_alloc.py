"""mose_chem_advance_cells, the chemistry entry point that allocates nothing, against mose_chem_advance.

The caller holds the cells in device memory: here plain hipMalloc buffers of the HIP runtime the library links
(libamdhip64), filled and read back with hipMemcpy.  Two calls in a row (the second on fewer cells, in a buffer of
the first call's stride) must give the states, hr, idid, step counts and report counters of mose_chem_advance on the
same cells bit for bit: both run the same kernels, only the work queue and statistics live in the library's static
scratch array instead of the chemistry object's buffer.  The refusals (too many cells, host memory) need no kernel.
"""
import ctypes as C

import numpy as np
import pytest

from mosegpu import Context, DeviceChem, MoseError, chem
from mosegpu.core import ThermoTables

pytestmark = pytest.mark.gpu

MECH = chem.h2_air_mechanism()
NZ = MECH.ns + 1
H2D, D2H = 1, 2


def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy):
        f.restype = C.c_int
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h, self.n, self.p = h, nbytes, C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.addr = self.p.value

    def put(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        a = np.zeros(shape, dtype=dtype)
        assert a.nbytes <= self.n
        assert self.h.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p.value:
            self.h.hipFree(self.p)
            self.p = C.c_void_p()


def allow_failed_cells(call, obj):
    """The report of an advance; cells that end with idid < 0 (RADAU5 stopping at NMAX) are allowed."""
    try:
        return call()
    except MoseError as e:
        assert "cells failed" in str(e)
        return obj.last_report


def reference(ctx, state, opts):
    dc = DeviceChem(ctx, MECH, state.shape[1])
    dc.upload(state)
    rep = allow_failed_cells(lambda: dc.advance(1e-6, opts), dc)
    out = dc.download()
    dc.destroy()
    return out, rep


def same_report(a, b):
    return all(getattr(a, k) == getattr(b, k) for k in ("cells", "failed", "steps_total", "rejected_total",
                                                         "rhs_total", "steps_max", "first_failed_idid",
                                                         "first_failed_cell"))


@pytest.mark.parametrize("method", [1, 2], ids=["RODAS", "RADAU5"])
def test_no_alloc_chemistry_equals_mose_chem_advance(gpu_ready, method):
    h = hip()
    ctx = Context(ThermoTables(MECH.wm, MECH.cp, MECH.h, names=MECH.names))
    owner = DeviceChem(ctx, MECH, 8)        # the mechanism; its own 8-cell buffer is not used
    stride = 240 if method == 1 else 72
    bufs = [DevBuf(h, NZ * stride * 8), DevBuf(h, stride * 8), DevBuf(h, stride * 4), DevBuf(h, stride * 4)]
    opts = chem.ode_opts(method=method)
    try:
        for call, n in enumerate((stride, stride // 3 + 1)):
            st = chem.h2_air_cells(n, seed=40 + call)
            full = np.zeros((NZ, stride))
            full[:, :n] = st
            bufs[0].put(full)
            print(f"no-alloc chemistry: method {method}, call {call}, {n} cells in a stride of {stride}", flush=True)
            rep = allow_failed_cells(lambda: owner.advance_cells(n, stride, bufs[0].addr, bufs[1].addr, bufs[2].addr,
                                                                 bufs[3].addr, 1e-6, opts), owner)
            got = bufs[0].get((NZ, stride), np.float64)[:, :n]
            hr = bufs[1].get(stride, np.float64)[:n]
            idid = bufs[2].get(stride, np.int32)[:n]
            nstep = bufs[3].get(stride, np.int32)[:n]
            (ref, hr_r, idid_r, nstep_r), rep_r = reference(ctx, st, opts)
            assert (idid == 1).mean() > 0.8
            assert np.array_equal(got, ref) and np.array_equal(hr, hr_r)
            assert np.array_equal(idid, idid_r) and np.array_equal(nstep, nstep_r)
            assert same_report(rep, rep_r) and rep.cells == n
    finally:
        for b in bufs:
            b.free()
        owner.destroy()
        ctx.close()


def test_no_alloc_chemistry_refusals(gpu_ready):
    h = hip()
    ctx = Context(ThermoTables(MECH.wm, MECH.cp, MECH.h, names=MECH.names))
    owner = DeviceChem(ctx, MECH, 8)
    b = [DevBuf(h, NZ * 16 * 8), DevBuf(h, 16 * 8), DevBuf(h, 16 * 4), DevBuf(h, 16 * 4)]
    host = np.zeros(NZ * 16)
    opts = chem.ode_opts()
    try:
        with pytest.raises(MoseError, match="ncell must be in"):
            owner.advance_cells(2 ** 31, 2 ** 31, b[0].addr, b[1].addr, b[2].addr, b[3].addr, 1e-6, opts)
        with pytest.raises(MoseError, match="stride < ncell"):
            owner.advance_cells(16, 8, b[0].addr, b[1].addr, b[2].addr, b[3].addr, 1e-6, opts)
        with pytest.raises(MoseError, match="device memory"):
            owner.advance_cells(16, 16, host.ctypes.data, b[1].addr, b[2].addr, b[3].addr, 1e-6, opts)
    finally:
        for x in b:
            x.free()
        owner.destroy()
        ctx.close()
