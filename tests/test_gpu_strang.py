"""Parity of the HIP path with the oracle on reacting hydrogen-air boxes: nine species (the residual kernels' nsc = 9
instantiation), chemistry on the case's resident state, Strang-split steps (half transport, Chemistry_Newstate_blk
over dt, half transport; Mod_Explicit.f90:155-294).

The bar for chemistry is a relative L-infinity <= 1e-8 per primitive variable after N steps; the device runs the
oracle's statements (the flux kernels' and the integrators' order, the controllers' glibc pow), so the tests also
report how many values differ bitwise.
"""
import numpy as np
import pytest

import mosegpu
from mosegpu import Context, DeviceCase, DeviceChem, chem, cases
from oracle_bind import OracleCase, OracleChem
from test_gpu_walls_sa import compare

pytestmark = pytest.mark.gpu

TOL = 1e-8
MECH = chem.h2_air_mechanism()


def run_reacting(case, nsteps, path, tile, method=1, attach=True):
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dch = DeviceChem(ctx, MECH, 1)
    dc.upload(case.P)
    chosen = dc.tune(path, tile)
    if attach:
        dc.chem_attach(dch, chem.ode_opts(method=method))
    rep = dc.step(nsteps)
    crep = dc.chem_report() if attach else None
    P = dc.download()
    dc.destroy()
    dch.destroy()
    ctx.close()
    return P, rep, crep, chosen


@pytest.mark.parametrize("viscous", [False, True], ids=["euler", "navier-stokes"])
@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (8, 4, 4)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_reacting_box_strang_steps(gpu_ready, viscous, path, tile):
    case = cases.reacting_box((12, 10, 8), viscous=viscous, T_range=(1200.0, 1700.0))
    nsteps = 3
    print(f"reacting box 12x10x8, viscous {viscous}, path {path}, tile {tile}, {nsteps} Strang steps", flush=True)
    Pg, rep, crep, chosen = run_reacting(case, nsteps, path, tile)
    o = OracleCase(case)
    o.set_chemistry(OracleChem(MECH, nthreads=16))
    Po = case.P.copy()
    dts = o.step(Po, nsteps)
    assert o.chem_failed() == 0 and crep.failed == 0
    assert crep.cells == 12 * 10 * 8 and crep.steps_total > 0
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    err = compare(case, Pg, Po, f"reacting box viscous {viscous} path {path} tile {chosen} x {nsteps} Strang steps")
    assert err <= TOL


def test_case_chemistry_advance_alone(gpu_ready):
    """mose_case_chem_advance: Chemistry_Newstate_blk on the resident state alone, against the oracle's."""
    case = cases.reacting_box((10, 8, 6), T_range=(1300.0, 1800.0))
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dch = DeviceChem(ctx, MECH, 1)
    dc.upload(case.P)
    dc.chem_attach(dch, chem.ode_opts())
    rep = dc.chem_advance(2e-6)
    Pg = dc.download()
    dc.destroy(); dch.destroy(); ctx.close()
    o = OracleCase(case)
    o.set_chemistry(OracleChem(MECH, nthreads=16))
    Po = case.P.copy()
    o.react(Po, 2e-6)
    assert rep.failed == 0 and rep.cells == 480
    assert compare(case, Pg, Po, "case chemistry alone, 10x8x6, dt 2e-6") <= TOL


def test_detached_chemistry_steps_are_transport_steps(gpu_ready):
    """Attach then detach: plain transport steps of the nine-species case, as the oracle's without chemistry."""
    case = cases.reacting_box((12, 10, 8))
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dch = DeviceChem(ctx, MECH, 1)
    dc.upload(case.P)
    dc.tune(mosegpu.PATH_FUSED, (0, 0, 0))
    dc.chem_attach(dch, chem.ode_opts())
    dc.chem_attach(None)
    dc.step(2)
    Pg = dc.download()
    dc.destroy(); dch.destroy(); ctx.close()
    Po = case.P.copy()
    OracleCase(case).step(Po, 2)
    assert compare(case, Pg, Po, "nine-species transport, 12x10x8 x 2") <= 1e-12
