"""Parity of the HIP path with the oracle on four-species cases (the residual kernels' nsc = 4 instantiation).

- The reference's own test/1D/diff1 inputs: CH4, H2O, N2, O2 layers at rest with Schmidt 1 diffusion, periodic
  faces in x, symmetry faces in y and no-flux faces in z (fixtures under tests/golden/cases).
- The four-copies viscous periodic box of the oracle's pins, in 3D.
The bar is a relative L-infinity <= 1e-12 per primitive variable; each test also reports how many values differ
bitwise.
"""
import numpy as np
import pytest

import mosegpu
from mosegpu import caseio, cases
from oracle_bind import OracleCase
from test_gpu_walls_sa import compare, run_device
from test_oracle_species import DIFF1, Y4, as_four_species

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (128, 1, 1)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_diff1_device_matches_oracle(gpu_ready, path, tile):
    case = caseio.load_case(DIFF1, 100)
    Pg, rep, chosen = run_device(case, 100, path, tile)
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, 100)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    assert compare(case, Pg, Po, f"diff1 (reference inputs, 4 species), path {path} tile {chosen} x 100") <= TOL


@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (8, 4, 4)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_four_species_box(gpu_ready, path, tile):
    c1 = cases.box_case((12, 10, 9), {f: "periodic" for f in range(1, 7)}, extent=(1.0, 0.9, 0.8), viscous=True,
                        mu_of_T=cases.sutherland(1.8e-3, 300.0), riemann="LMRoe", limiter="vanleer")
    rng = np.random.Generator(np.random.PCG64(13))
    V = c1.layout.interior(c1.P, 0)
    V[0] = 1.1 + 0.1 * rng.uniform(-1, 1, V[0].shape)
    for q in range(3):
        V[1 + q] = 30.0 * rng.uniform(-1, 1, V[0].shape)
    V[4] = 1e5 * (1.0 + 0.05 * rng.uniform(-1, 1, V[0].shape))
    cases.fill_ghosts_nearest(c1.layout, c1.P)
    Yr = Y4 * (1.0 + 0.2 * rng.uniform(-1, 1, 4))          # uneven fractions: the species fluxes differ
    case = as_four_species(c1, Yr / Yr.sum(), "LMRoe")
    W = case.layout.interior(case.P, 0)
    W[0] *= 1.0 + 0.3 * rng.uniform(-1, 1, W[0].shape)     # and a composition that varies in space
    cases.fill_ghosts_nearest(case.layout, case.P)
    Pg, rep, chosen = run_device(case, 4, path, tile)
    Po = case.P.copy()
    OracleCase(case).step(Po, 4)
    assert compare(case, Pg, Po, f"four-species viscous box, path {path} tile {chosen} x 4") <= TOL
