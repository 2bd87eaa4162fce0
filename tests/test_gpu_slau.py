"""Parity of the HIP path with the oracle for the SLAU flux (riemann-solver = SLAU: the interior kernels'
conv_physics.inc:126-167 and the record kernels' riemann_SLAU, src/lib/numerics/fluxes/Lib_Riemann.f90:140-198).

The reference's own test/1D/Sod79 inputs (SLAU + MC limiter, CFL 0.9) and a periodic Taylor-Green box (every boundary
face a connection record, so the records' SLAU runs too), on the parity and plain paths: equal to the oracle bit for
bit.  The performance path refuses SLAU."""
import os

import numpy as np
import pytest

import mosegpu
from mosegpu import Context, DeviceCase, MoseError, caseio, cases
from oracle_bind import OracleCase
from test_gpu_walls_sa import compare, interiors, run_device

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")
FUSED, PLAIN = mosegpu.PATH_FUSED, mosegpu.PATH_PLAIN


@pytest.mark.parametrize("path,tile", [(FUSED, (0, 0, 0)), (FUSED, (128, 1, 1)), (PLAIN, (0, 0, 0))])
def test_sod79_reference_inputs(gpu_ready, path, tile):
    c = caseio.load_case(os.path.join(GOLD, "sod79"), 100)
    assert c.scheme.riemann == 3
    print(f"Sod79 (reference inputs, SLAU + MC), path {path}, tile {tile}, 100 steps", flush=True)
    Pg, rep, chosen = run_device(c, 100, path, tile)
    Po = c.P.copy()
    dts = OracleCase(c).step(Po, 100)
    compare(c, Pg, Po, f"Sod79 SLAU path {path} tile {chosen} x 100")
    assert abs(rep.time - dts.sum()) <= 1e-15 * dts.sum()
    assert np.array_equal(interiors(c, Pg), interiors(c, Po))


@pytest.mark.parametrize("path,tile", [(FUSED, (8, 8, 4)), (PLAIN, (0, 0, 0))])
def test_slau_periodic_taylor_green(gpu_ready, path, tile):
    c = cases.taylor_green(16, 6)
    c.scheme.riemann = 3
    print(f"Taylor-Green 16^3 with SLAU, path {path}, tile {tile}, 6 steps", flush=True)
    Pg, rep, chosen = run_device(c, 6, path, tile)
    Po = c.P.copy()
    OracleCase(c, nthreads=16).step(Po, 6)
    compare(c, Pg, Po, f"Taylor-Green 16^3 SLAU path {path} tile {chosen} x 6")
    assert np.array_equal(interiors(c, Pg), interiors(c, Po))


def test_slau_is_refused_by_the_performance_path(gpu_ready):
    c = caseio.load_case(os.path.join(GOLD, "sod79"), 1)
    ctx = Context(c.tables)
    dc = DeviceCase(ctx, c)
    with pytest.raises(MoseError):
        dc.tune(mosegpu.PATH_FAST, (0, 0, 0))
    dc.destroy()
    ctx.close()
