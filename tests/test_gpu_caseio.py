"""The reference's own isentropic-vortex case (its input.ini, bc.txt, ic.tec, phase.txt and thermo.dat, as fixtures
under tests/golden/cases), read by mosegpu.caseio and run through the C ABI: the HIP path against the oracle.

The bar is a relative L-infinity <= 1e-12 per primitive variable after 20 steps; both sides keep the reference's
rounding order, and each comparison reports how many values differ bitwise.  The periodic box also pins the device
result on its own: mass and momentum are conserved to round-off.
"""
import os

import numpy as np
import pytest

import mosegpu
from mosegpu import Context, DeviceCase, caseio
from oracle_bind import OracleCase

pytestmark = pytest.mark.gpu

TOL = 1e-12
VORTEX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "isentropic-vortex")
NSTEPS = 20


@pytest.fixture(scope="module")
def vortex_oracle():
    c = caseio.load_case(VORTEX, nsteps=NSTEPS)
    Po = c.P.copy()
    dts = OracleCase(c, nthreads=16).step(Po, NSTEPS)
    return c, Po, dts


@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (16, 16, 1)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_vortex_device_matches_oracle(gpu_ready, vortex_oracle, path, tile):
    c, Po, dts = vortex_oracle
    ctx = Context(c.tables)
    dc = DeviceCase(ctx, c)
    dc.upload(c.P)
    chosen = dc.tune(path, tile)
    rep = dc.step(NSTEPS)
    Pg = dc.download()
    dc.destroy()
    ctx.close()
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    a = c.layout.interior(Pg, 0).reshape(c.nprim, -1)
    b = c.layout.interior(Po, 0).reshape(c.nprim, -1)
    scale = np.abs(b).max(axis=1)
    scale[scale == 0] = 1.0
    err = (np.abs(a - b).max(axis=1) / scale).max()
    print(f"isentropic vortex (reference inputs), path {path} tile {tuple(chosen)} x {NSTEPS}: rel Linf {err:.3e}, "
          f"values differing bitwise {int(np.sum(a != b))} of {a.size}", flush=True)
    assert err <= TOL
    vol = c.layout.interior(c.geom.vol[None, :], 0)[0].reshape(-1)

    def totals(P):
        V = c.layout.interior(P, 0).reshape(c.nprim, -1)
        return np.array([np.sum(V[0] * vol), np.sum(V[0] * V[1] * vol), np.sum(V[0] * V[2] * vol)])

    t0, t1 = totals(c.P), totals(Pg)
    assert np.all(np.abs(t1 - t0) <= 1e-13 * np.abs(t0))


@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (8, 8, 4)), (mosegpu.PATH_FUSED, (8, 8, 8)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_none_faces_viscous_box(gpu_ready, path, tile):
    """MOSE 0 faces in a viscous 3D box (periodic x, walls y, no-flux z): their zero-gradient ghosts enter the
    tangential gradients of the viscous faces next to them."""
    from test_caseio import none_face_box
    c = none_face_box()
    Po = c.P.copy()
    dts = OracleCase(c, nthreads=16).step(Po, 5)
    ctx = Context(c.tables)
    dc = DeviceCase(ctx, c)
    dc.upload(c.P)
    chosen = dc.tune(path, tile)
    rep = dc.step(5)
    Pg = dc.download()
    dc.destroy()
    ctx.close()
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    a = c.layout.interior(Pg, 0).reshape(c.nprim, -1)
    b = c.layout.interior(Po, 0).reshape(c.nprim, -1)
    scale = np.abs(b).max(axis=1)
    scale[scale == 0] = 1.0
    err = (np.abs(a - b).max(axis=1) / scale).max()
    print(f"MOSE 0 faces, viscous box, path {path} tile {tuple(chosen)} x 5: rel Linf {err:.3e}, "
          f"values differing bitwise {int(np.sum(a != b))} of {a.size}", flush=True)
    assert err <= TOL
