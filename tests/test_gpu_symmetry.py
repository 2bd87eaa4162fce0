"""Parity of the HIP path with the oracle on boxes with symmetry faces (MOSE 300, internal type 3).

The cases are the half boxes of the oracle's mirror pins: symmetry faces on one axis, periodic faces on the others
(walls on the y faces for Spalart-Allmaras), a seeded state, and NaN in the second ghost layer behind the symmetry
faces, which no stencil may read.  The bar is a relative L-infinity <= 1e-12 per primitive variable after 4 steps;
each test also reports how many values differ bitwise.
"""
import numpy as np
import pytest

import mosegpu
from oracle_bind import OracleCase
from test_gpu_walls_sa import compare, run_device
from test_oracle_symmetry import mirror_pair

pytestmark = pytest.mark.gpu

TOL = 1e-12
PATHS = [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (8, 8, 4)), (mosegpu.PATH_FUSED, (8, 4, 4)),
         (mosegpu.PATH_PLAIN, (0, 0, 0))]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("viscous", [False, True])
@pytest.mark.parametrize("path,tile", PATHS)
def test_symmetry_box(gpu_ready, axis, viscous, path, tile):
    case, _ = mirror_pair(axis, [12, 10, 9], viscous, riemann="HLLC" if viscous else "LMRoe", seed=31 + axis)
    Pg, rep, chosen = run_device(case, 4, path, tile)
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, 4)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    err = compare(case, Pg, Po, f"symmetry axis {axis} viscous {viscous} path {path} tile {chosen} x 4")
    assert err <= TOL


@pytest.mark.parametrize("path,tile", PATHS)
@pytest.mark.parametrize("viscous", [False, True])
def test_symmetry_box_hlle(gpu_ready, path, tile, viscous):
    """HLLE (rs 2) in the interior faces and in BC_Symmetry_Eul's face flux."""
    case, _ = mirror_pair(1, [12, 10, 9], viscous, riemann="HLLE", seed=51)
    Pg, rep, chosen = run_device(case, 4, path, tile)
    Po = case.P.copy()
    OracleCase(case).step(Po, 4)
    assert compare(case, Pg, Po, f"symmetry box, HLLE, viscous {viscous}, path {path} tile {chosen} x 4") <= TOL


@pytest.mark.parametrize("path,tile", PATHS)
def test_symmetry_box_sa(gpu_ready, path, tile):
    case, _ = mirror_pair(0, [12, 10, 9], True, nrans=1, walls=1, seed=41)
    Pg, rep, chosen = run_device(case, 4, path, tile)
    Po = case.P.copy()
    OracleCase(case).step(Po, 4)
    assert compare(case, Pg, Po, f"symmetry + SA path {path} tile {chosen} x 4") <= TOL
