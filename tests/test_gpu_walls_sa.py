"""Parity of the HIP path with the oracle on multi-block lattices with walls, far field and Spalart-Allmaras.

The bar is a relative L-infinity <= 1e-12 per primitive variable after N steps. Both sides keep the reference's
rounding order without FMA contraction, and each test also reports how many values differ bitwise.
"""
import numpy as np
import pytest

import mosegpu
from mosegpu import Context, DeviceCase, cases
from mosegpu.core import BC_FARFIELD, BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL, FARFIELD_DTYPE
from oracle_bind import OracleCase

pytestmark = pytest.mark.gpu

TOL = 1e-12


def interiors(case, P):
    return np.hstack([case.layout.interior(P, b).reshape(case.nprim, -1) for b in range(case.layout.nb)])


def compare(case, Pg, Po, what):
    a, b = interiors(case, Pg), interiors(case, Po)
    scale = np.abs(b).max(axis=1)
    scale[scale == 0] = 1.0
    err = (np.abs(a - b).max(axis=1) / scale).max()
    nbit = int(np.sum(a != b))
    print(f"{what}: rel Linf {err:.3e}, values differing bitwise {nbit} of {a.size}", flush=True)
    return err


def run_device(case, nsteps, path=mosegpu.PATH_FUSED, tile=(0, 0, 0)):
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    chosen = dc.tune(path, tile)
    rep = dc.step(nsteps)
    P = dc.download()
    dc.destroy()
    ctx.close()
    return P, rep, chosen


@pytest.mark.parametrize("wall", [BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL])
@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (16, 4, 4)),
                                       (mosegpu.PATH_FUSED, (8, 4, 4)), (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_channel_lattice_small(gpu_ready, wall, path, tile):
    case = cases.channel_lattice(nb=(2, 2, 2), wmin=4, wmax=9, seed=11, nsteps=5, wall=wall, T_wall=280.0)
    Pg, rep, chosen = run_device(case, 5, path, tile)
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, 5)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    assert compare(case, Pg, Po, f"channel 2x2x2 wall {wall} path {path} tile {chosen} x 5") <= TOL


def test_channel_lattice_4x4x4(gpu_ready):
    case = cases.channel_lattice(nb=(4, 4, 4), wmin=6, wmax=12, seed=5, nsteps=3, riemann="LMRoe")
    Pg, rep, chosen = run_device(case, 3)
    Po = case.P.copy()
    OracleCase(case, nthreads=16).step(Po, 3)
    assert compare(case, Pg, Po, f"channel 4x4x4 blocks ({case.interior_cells()} cells), tile {chosen} x 3") <= TOL


@pytest.mark.parametrize("rs", ["HLLC", "AUSM+", "LMRoe"])
def test_farfield_wall_residual(gpu_ready, rs):
    """Residual parity on a perturbed state in a box with far field on five faces and a wall on one: inflow and
    outflow faces, subsonic and closed far-field faces all occur."""
    rho, p, M = 1.15, 9.0e4, 0.4
    T = p / (rho * cases.R_AIR)
    gam = cases.CP_AIR / (cases.CP_AIR - cases.R_AIR)
    U = M * np.sqrt(gam * cases.R_AIR * T)
    T0 = T * (1.0 + 0.5 * (gam - 1.0) * M * M)
    ff = np.zeros(1, dtype=FARFIELD_DTYPE)
    ff["t0"] = T0; ff["p0"] = p * (T0 / T) ** (gam / (gam - 1.0)); ff["pamb"] = p; ff["rel_fac"] = 1.0
    ff["ci"][0, 0] = 1.0; ff["rans"][0, 0] = 1e-4
    faces = {1: BC_FARFIELD, 2: BC_FARFIELD, 3: BC_WALL_ADIABATIC, 4: BC_FARFIELD, 5: BC_FARFIELD, 6: BC_FARFIELD}
    case = cases.box_case((9, 7, 6), faces, extent=(1.0, 0.6, 0.5), farfield=ff, nrans=1, riemann=rs, point_implicit=0)
    rng = np.random.Generator(np.random.PCG64(17))
    V = case.layout.interior(case.P, 0)
    sh = V.shape[1:]
    V[0] = rho * (1 + 0.05 * rng.uniform(-1, 1, sh))
    V[1] = U * (1 + 0.3 * rng.uniform(-1, 1, sh))
    V[2] = 0.2 * U * rng.uniform(-1, 1, sh)
    V[3] = 0.2 * U * rng.uniform(-1, 1, sh)
    V[4] = p * (1 + 0.05 * rng.uniform(-1, 1, sh))
    V[5] = 1.8e-5 * rng.uniform(1, 5, sh)
    cases.fill_ghosts_nearest(case.layout, case.P)
    yc = (np.arange(7) + 0.5) * (0.6 / 7)
    case.layout.interior(case.geom.yn, 0)[...] = yc[None, :, None]
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    Rg = dc.residual()
    Ro = OracleCase(case).residual(case.P.copy(), 0.0)
    assert compare(case, Rg, Ro, f"far field + wall residual ({rs})") <= TOL


@pytest.mark.parametrize("rs", ["HLLC", "LMRoe", "AUSM+"])
def test_walls_farfield_steps_on_512_cell_tiles(gpu_ready, rs):
    """Five steps of a viscous box with far field on five faces and an isothermal wall on one, without RANS (a
    512-cell tile's LDS holds five primitives, not six), on the 8 x 8 x 8 tiles of 512-thread workgroups."""
    rho, p, M = 1.15, 9.0e4, 0.3
    T = p / (rho * cases.R_AIR)
    gam = cases.CP_AIR / (cases.CP_AIR - cases.R_AIR)
    U = M * np.sqrt(gam * cases.R_AIR * T)
    T0 = T * (1.0 + 0.5 * (gam - 1.0) * M * M)
    ff = np.zeros(1, dtype=FARFIELD_DTYPE)
    ff["t0"] = T0; ff["p0"] = p * (T0 / T) ** (gam / (gam - 1.0)); ff["pamb"] = p; ff["rel_fac"] = 1.0
    ff["ci"][0, 0] = 1.0
    faces = {1: BC_FARFIELD, 2: BC_FARFIELD, 3: BC_WALL_ISOTHERMAL, 4: BC_FARFIELD, 5: BC_FARFIELD, 6: BC_FARFIELD}
    case = cases.box_case((19, 13, 11), faces, extent=(1.0, 0.6, 0.5), farfield=ff, riemann=rs, T_wall=290.0,
                          limiter="vanalbada", cfl=0.4)
    rng = np.random.Generator(np.random.PCG64(23))
    V = case.layout.interior(case.P, 0)
    sh = V.shape[1:]
    V[0] = rho * (1 + 0.02 * rng.uniform(-1, 1, sh))
    V[1] = U * (1 + 0.1 * rng.uniform(-1, 1, sh))
    V[2] = 0.05 * U * rng.uniform(-1, 1, sh)
    V[3] = 0.05 * U * rng.uniform(-1, 1, sh)
    V[4] = p * (1 + 0.02 * rng.uniform(-1, 1, sh))
    cases.fill_ghosts_nearest(case.layout, case.P)
    Pg, rep, chosen = run_device(case, 5, mosegpu.PATH_FUSED, (8, 8, 8))
    assert tuple(chosen) == (8, 8, 8)
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, 5)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    assert compare(case, Pg, Po, f"walls + far field, {rs}, 512-cell tiles x 5") <= TOL

