"""Parity of the HIP path with the oracle on the reference's supersonic forward step inputs
(test/2D/euler/supersonic-forward-step: two blocks joined by connection records, MOSE_open 405 supersonic inlets
through BC_Inflow_Outflow's supersonic closure, Lib_BC.f90:1574-1583 and 1247-1330, symmetry walls, an extrapolation
exit; HLLE + van Leer, CFL 0.8).  The parity and plain paths equal the oracle bit for bit; the performance path is
held to a relative L-infinity of 1e-12."""
import os

import numpy as np
import pytest

import mosegpu
from mosegpu import caseio
from oracle_bind import OracleCase
from test_gpu_fast_path import fast_error
from test_gpu_walls_sa import compare, interiors, run_device

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "forward-step")
FUSED, PLAIN, FAST = mosegpu.PATH_FUSED, mosegpu.PATH_PLAIN, mosegpu.PATH_FAST


@pytest.mark.parametrize("path,tile", [(FUSED, (0, 0, 0)), (FUSED, (16, 8, 1)), (PLAIN, (0, 0, 0))])
def test_forward_step_reference_inputs(gpu_ready, path, tile):
    c = caseio.load_case(GOLD, 200)
    print(f"supersonic forward step (reference inputs), path {path}, tile {tile}, 200 steps", flush=True)
    Pg, rep, chosen = run_device(c, 200, path, tile)
    Po = c.P.copy()
    dts = OracleCase(c, nthreads=16).step(Po, 200)
    compare(c, Pg, Po, f"forward step path {path} tile {chosen} x 200")
    assert abs(rep.time - dts.sum()) <= 1e-15 * dts.sum()
    assert np.array_equal(interiors(c, Pg), interiors(c, Po))


def test_forward_step_performance_path(gpu_ready):
    c = caseio.load_case(GOLD, 200)
    print("supersonic forward step (reference inputs), performance path, 200 steps", flush=True)
    Pg, rep, chosen = run_device(c, 200, FAST, (0, 0, 0))
    Po = c.P.copy()
    OracleCase(c, nthreads=16).step(Po, 200)
    assert fast_error(c, Pg, Po, f"forward step fast path tile {chosen} x 200") <= 1e-12
