"""How far a change of rounding alone moves the reference's own test cases: the oracle compiled with fused
multiply-add contraction (oracle/Makefile, target fma) against the oracle as built (no contraction), both on the
reference's inputs for the same number of steps.

The performance path (MOSE_PATH_FAST) differs from the parity path by rounding only, so these distances are the floor
of what any non-bitwise build can promise on each case -- the reference's own Fortran compiled with -O3 -march=native
(which contracts) included:
  - the isentropic vortex, Shu-Osher and Ein91 move by a few 1e-15 .. 1e-14 (relative L-infinity per variable, the
    velocity in units of |u| + sqrt(p / rho)); their times not at all or by about 1e-16;
  - Noh87 moves by about 2e-11, its time by about 2e-12: first-order HLLE on a cold gas whose temperature comes from
    the total energy minus a kinetic energy about a million times larger, so each rounding of the energy becomes a
    relative change of order 1e-10 of the cold gas's temperature and pressure, and through the CFL limit of the
    shocked region a change of the dt sequence.
The GPU tests hold the performance path to 1e-12 on every case whose floor here is below 1e-13, and to 1e-9 on Noh87.
"""
import os
import pathlib
import subprocess

import numpy as np
import pytest

import oracle_bind
from mosegpu import caseio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")
ORACLE_DIR = pathlib.Path(oracle_bind.ORACLE_DIR)


def _rel(case, a, b):
    ns = case.scheme.nsc
    A = np.hstack([case.layout.interior(a, q).reshape(case.nprim, -1) for q in range(case.layout.nb)])
    B = np.hstack([case.layout.interior(b, q).reshape(case.nprim, -1) for q in range(case.layout.nb)])
    scale = np.abs(B).max(axis=1)
    rho = B[:ns].sum(axis=0)
    scale[ns:ns + 3] = float((np.sqrt((B[ns:ns + 3] ** 2).sum(axis=0)) + np.sqrt(B[ns + 3] / rho)).max())
    scale[scale == 0] = 1.0
    return float((np.abs(A - B).max(axis=1) / scale).max())


@pytest.fixture(scope="module")
def contracted():
    subprocess.run(["make", "-C", str(ORACLE_DIR), "fma"], check=True, capture_output=True)
    return str(ORACLE_DIR / "libmoseoracle_fma.so")


def _run(lib, case, n):
    """n steps of the oracle library `lib` (a separate load of the same API) from the case's start."""
    saved = (oracle_bind.ORACLE_LIB, oracle_bind._L)
    try:
        oracle_bind.ORACLE_LIB = pathlib.Path(lib)
        oracle_bind._L = None
        P = case.P.copy()
        dts = oracle_bind.OracleCase(case, nthreads=8).step(P, n)
    finally:
        oracle_bind.ORACLE_LIB, oracle_bind._L = saved
    return P, float(dts.sum())


@pytest.mark.parametrize("name,n,floor_max,floor_min", [
    ("isentropic-vortex", 20, 1e-13, 0.0),
    ("shu-osher", 330, 1e-13, 0.0),
    ("ein91", 100, 1e-13, 0.0),
    ("noh87", 100, 1e-9, 1e-12),
])
def test_rounding_floor_of_the_reference_cases(contracted, name, n, floor_max, floor_min):
    case = caseio.load_case(os.path.join(GOLD, name), n)
    Pa, ta = _run(contracted, case, n)
    Pb, tb = _run(str(oracle_bind.ORACLE_LIB), case, n)
    err, dt = _rel(case, Pa, Pb), abs(ta - tb) / tb
    print(f"{name} x {n}: contracted oracle vs oracle: rel Linf {err:.3e}, time {dt:.3e}", flush=True)
    assert err <= floor_max and dt <= floor_max
    assert err >= floor_min
