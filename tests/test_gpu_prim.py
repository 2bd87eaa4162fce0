"""Parity of the HIP path with the oracle for the primitive-variable update (integration-variables = prim,
Newstate_Primitive_blk, src/lib/numerics/time/explicit/Lib_Newstate.f90:340-531).

The parity and plain paths run the oracle's statements and must equal it bit for bit: one species (Sod, the viscous
Taylor-Green box), several species (the mass-fraction sensitivities of the update), Spalart-Allmaras on a walled
channel (the turbulence variable) and the nine-species reacting box with Strang-split chemistry.  The performance path
is held to its bar, a relative L-infinity of 1e-12, below nine species; its nine-species kernels refuse the update.
"""
import numpy as np
import pytest

import mosegpu
from mosegpu import Context, DeviceCase, DeviceChem, MoseError, cases, chem
from oracle_bind import OracleCase, OracleChem
from test_gpu_fast_path import fast_error
from test_gpu_walls_sa import compare, interiors, run_device
from test_oracle_species import Y4, as_four_species

pytestmark = pytest.mark.gpu

FUSED, PLAIN, FAST = mosegpu.PATH_FUSED, mosegpu.PATH_PLAIN, mosegpu.PATH_FAST


def prim(case):
    case.scheme.integration = mosegpu.INT_PRIM
    return case


def parity(case, nsteps, path, tile, what, bitwise=True):
    """bitwise: the device's and the oracle's arithmetic are the same statements; cases with far-field faces or the SA
    source are held to 1e-12 instead, since their real powers come from the device's and the host's pow"""
    print(f"{what}: path {path}, tile {tile}, {nsteps} steps", flush=True)
    Pg, rep, chosen = run_device(case, nsteps, path, tile)
    Po = case.P.copy()
    dts = OracleCase(case, nthreads=16).step(Po, nsteps)
    upd = "primitive" if case.scheme.integration == mosegpu.INT_PRIM else "conservative"
    err = compare(case, Pg, Po, f"{what} ({upd} update) path {path} tile {chosen} x {nsteps}")
    assert abs(rep.time - dts.sum()) <= (1e-15 if bitwise else 1e-13) * dts.sum()
    if bitwise:
        assert np.array_equal(interiors(case, Pg), interiors(case, Po))
    assert err <= 1e-12
    return Pg, Po


def moving_box(nsc, seed=7):
    c1 = cases.box_case((10, 9, 8), {f: "periodic" for f in range(1, 7)}, extent=(1.0, 0.9, 0.8), viscous=True,
                        mu_of_T=cases.sutherland(1.8e-3, 300.0), riemann="LMRoe", limiter="vanalbada")
    rng = np.random.Generator(np.random.PCG64(seed))
    V = c1.layout.interior(c1.P, 0)
    V[0] = 1.1 + 0.1 * rng.uniform(-1, 1, V[0].shape)
    for q in range(3):
        V[1 + q] = 30.0 * rng.uniform(-1, 1, V[0].shape)
    V[4] = 1e5 * (1.0 + 0.05 * rng.uniform(-1, 1, V[0].shape))
    cases.fill_ghosts_nearest(c1.layout, c1.P)
    if nsc == 1:
        return prim(c1)
    c4 = as_four_species(c1, Y4, "LMRoe")
    # four species of different molar masses and formation enthalpies, so that drho/dY and dh/dY of the update are
    # not zero
    c4.tables.wm[:] = [28.0, 32.0, 18.0, 40.0]
    c4.tables.h += np.array([0.0, 3.0e4, -1.0e5, 5.0e4])[:, None]
    return prim(c4)


@pytest.mark.parametrize("path,tile", [(FUSED, (8, 8, 4)), (FUSED, (16, 4, 2)), (PLAIN, (0, 0, 0))])
def test_prim_sod(gpu_ready, path, tile):
    tile = (32, 4, 2) if path == FUSED and tile == (8, 8, 4) else tile
    parity(prim(cases.sod(400, 60)), 60, path, tile, "Sod 400")


@pytest.mark.parametrize("path,tile", [(FUSED, (8, 8, 4)), (FUSED, (8, 8, 8)), (PLAIN, (0, 0, 0))])
def test_prim_taylor_green(gpu_ready, path, tile):
    parity(prim(cases.taylor_green(16, 8)), 8, path, tile, "Taylor-Green 16^3")


@pytest.mark.parametrize("nsc", [1, 4])
@pytest.mark.parametrize("path,tile", [(FUSED, (16, 4, 2)), (PLAIN, (0, 0, 0))])
def test_prim_moving_box(gpu_ready, nsc, path, tile):
    parity(moving_box(nsc), 6, path, tile, f"moving viscous box, {nsc} species")


def test_prim_channel_sa(gpu_ready):
    case = prim(cases.channel_lattice(nb=(2, 2, 2), wmin=4, wmax=9, seed=11, nsteps=5, riemann="LMRoe"))
    assert case.scheme.nrans == 1
    parity(case, 5, FUSED, (8, 4, 4), "SA channel 2x2x2", bitwise=False)


def test_prim_reacting_box(gpu_ready):
    mech = chem.h2_air_mechanism()
    case = prim(cases.reacting_box((12, 10, 8), viscous=True, T_range=(1200.0, 1700.0)))
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dch = DeviceChem(ctx, mech, 1)
    dc.upload(case.P)
    dc.tune(FUSED, (16, 4, 2))
    dc.chem_attach(dch, chem.ode_opts())
    print("reacting box 12x10x8 (primitive update), 2 Strang steps", flush=True)
    rep = dc.step(2)
    Pg = dc.download()
    with pytest.raises(MoseError):      # the performance path's nine-species kernels: conservative update only
        dc.tune(FAST, (0, 0, 0))
    dc.destroy(); dch.destroy(); ctx.close()
    o = OracleCase(case, nthreads=16)
    o.set_chemistry(OracleChem(mech, nthreads=16))
    Po = case.P.copy()
    dts = o.step(Po, 2)
    assert abs(rep.time - dts.sum()) <= 1e-15 * dts.sum()
    compare(case, Pg, Po, "reacting box (primitive update) x 2 Strang steps")
    assert np.array_equal(interiors(case, Pg), interiors(case, Po))


@pytest.mark.parametrize("what", ["sod", "tgv", "four"])
def test_prim_fast_path(gpu_ready, what):
    case = {"sod": lambda: prim(cases.sod(400, 60)), "tgv": lambda: prim(cases.taylor_green(24, 8)),
            "four": lambda: moving_box(4)}[what]()
    nsteps = {"sod": 60, "tgv": 8, "four": 6}[what]
    tile = {"sod": (32, 4, 2), "tgv": (8, 8, 4), "four": (16, 4, 2)}[what]
    Pg, rep, chosen = run_device(case, nsteps, FAST, tile)
    Po = case.P.copy()
    OracleCase(case, nthreads=16).step(Po, nsteps)
    assert fast_error(case, Pg, Po, f"fast path, primitive update: {what} tile {chosen} x {nsteps}") <= 1e-12


def test_prim_differs_from_cons(gpu_ready):
    """The flag reaches the kernels: the same case with the conservative update ends elsewhere (by O(dt^2))."""
    a = cases.taylor_green(16, 4)
    b = prim(cases.taylor_green(16, 4))
    Pa, _, _ = run_device(a, 4, FUSED, (8, 8, 4))
    Pb, _, _ = run_device(b, 4, FUSED, (8, 8, 4))
    d = np.abs(interiors(a, Pa) - interiors(b, Pb)).max(axis=1) / np.abs(interiors(a, Pa)).max(axis=1)
    print(f"Taylor-Green 16^3 x 4, primitive against conservative update: relative difference per variable {d}")
    assert d.max() > 1e-12
