"""Host-side checks that need no GPU: the C-ABI library loads and exports every function include/mose_gpu.h
declares, the record / far-field structs match the header's sizes, the host geometry builder, and the host
restatement of field_state_hash."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import mosegpu
from mosegpu import cases
from mosegpu.core import FARFIELD_DTYPE, LIB_PATH, RECORD_DTYPE

HEADER = Path(__file__).resolve().parent.parent / "include" / "mose_gpu.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mose_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(str(LIB_PATH))
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_zero_or_more():
    n = C.c_int(-1)
    assert mosegpu.lib().mose_device_count(C.byref(n)) == 0
    assert n.value >= 0


def test_struct_sizes_match_header():
    assert RECORD_DTYPE.itemsize == 104
    assert FARFIELD_DTYPE.itemsize == 208
    assert RECORD_DTYPE.fields["tw"][1] == 72 and RECORD_DTYPE.fields["tvec"][1] == 80


def test_state_hash_of_known_bits():
    lay = mosegpu.Layout([(2, 1, 1)])
    P = np.zeros((5, lay.NC))
    V = lay.interior(P, 0)
    V[0] = 1.0          # 0x3FF0000000000000
    V[4] = -2.5         # 0xC004000000000000
    V[1, 0, 0, 1] = np.frombuffer(np.uint64(0x0000000100000002).tobytes(), dtype=np.float64)[0]
    lo, hi = mosegpu.state_hash(lay, P)
    assert lo == 2
    assert hi == 2 * 0x3FF00000 + 2 * 0xC0040000 + 1


def test_records_tangent_is_the_face_diagonal():
    c = cases.box_case((3, 4, 2), {f: mosegpu.BC_FARFIELD for f in range(1, 7)}, extent=(0.3, 0.8, 0.1),
                       viscous=False, farfield=np.zeros(1, dtype=FARFIELD_DTYPE))
    r = c.records
    hx, hy, hz = 0.1, 0.2, 0.05
    for f, t in ((1, (0, hy, hz)), (3, (hx, 0, hz)), (5, (hx, hy, 0))):
        tv = np.array(t) / np.linalg.norm(t)
        got = r[r["f"] == f]["tvec"]
        assert np.allclose(got, tv[None, :], rtol=1e-14, atol=1e-15)


def test_case_create_rejects_bad_records_without_gpu():
    # without a device the context cannot be made; plan errors are reported by mose_case_bytes' sizing instead
    b = C.c_int64(0)
    s = mosegpu.Scheme()
    s.nsc = 1
    dims = np.array([4, 4, 4], dtype=np.int32)
    assert mosegpu.lib().mose_case_bytes(C.byref(s), 1, dims.ctypes.data_as(C.POINTER(C.c_int)), 96, C.byref(b)) == 0
    assert b.value > 3 * 8 * 8 * 8 * 5 * 8
