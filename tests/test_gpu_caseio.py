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


SHU_OSHER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "shu-osher")


@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (128, 1, 1)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_shu_osher_device_matches_oracle(gpu_ready, path, tile):
    """The reference's test/1D/Shu-Osher inputs, 330 steps (to t ~ 1.77, its threshold being 1.8)."""
    c = caseio.load_case(SHU_OSHER, 330)
    ctx = Context(c.tables)
    dc = DeviceCase(ctx, c)
    dc.upload(c.P)
    chosen = dc.tune(path, tile)
    rep = dc.step(330)
    Pg = dc.download()
    dc.destroy()
    ctx.close()
    Po = c.P.copy()
    dts = OracleCase(c).step(Po, 330)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    a, b = c.layout.interior(Pg, 0), c.layout.interior(Po, 0)
    err = max(np.abs(a[s] - b[s]).max() / max(np.abs(b[s]).max(), 1.0) for s in range(c.nprim))
    print(f"shu-osher (reference inputs), path {path} tile {chosen} x 330: rel Linf {err:.3e}, values differing "
          f"bitwise {int(np.sum(a != b))} of {a.size}", flush=True)
    assert err <= TOL


def test_residual_history_file(gpu_ready, tmp_path):
    """caseio.step_with_history on the Shu-Osher inputs with diter-res = 7: rows at iterations 7, 14, 21, 28, then a
    restart that appends rows 35 and 42.  Each row's time is the oracle's, and its norms are the L2 norms of the
    oracle's change over that step (summed in another order, so to 1e-10)."""
    from test_gpu_parity import oracle_residual_norm
    c = caseio.load_case(SHU_OSHER, 45)
    ctx = Context(c.tables)
    dc = DeviceCase(ctx, c)
    dc.upload(c.P)
    dc.tune(mosegpu.PATH_FUSED, (0, 0, 0))
    path = str(tmp_path / "residual-history.dat")
    it, t = caseio.step_with_history(dc, 30, 7, path)
    assert it == 30
    it, t = caseio.step_with_history(dc, 15, 7, path, newrun=False)
    dc.destroy()
    ctx.close()
    assert it == 45
    rows = open(path).read().splitlines()
    assert [int(r[:8]) for r in rows] == [7, 14, 21, 28, 35, 42]
    assert all(len(r) == 8 + 20 * 6 for r in rows)
    o = OracleCase(c)
    P = c.P.copy()
    time = 0.0
    for n in range(1, 43):
        Pold = P.copy()
        time += o.step(P, 1)[0]
        if n % 7 == 0:
            r = rows[n // 7 - 1]
            vals = [float(r[8 + 20 * q:8 + 20 * (q + 1)].replace("E", "e")) for q in range(6)]
            assert abs(vals[0] - time) <= 1e-9 * time
            want = oracle_residual_norm(c, P, Pold)
            assert np.all(np.abs(np.array(vals[1:]) - want) <= 1e-9 * want.max() + 1e-300), (n, vals, want)
    print(f"residual history: {len(rows)} rows, last {rows[-1]}", flush=True)


EIN91 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "ein91")


@pytest.mark.parametrize("name", ["ein91", "noh87"])
@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (128, 1, 1)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_hlle_cases_device_matches_oracle(gpu_ready, name, path, tile):
    """The reference's test/1D/Ein91 and Noh87 inputs (HLLE), 100 steps."""
    c = caseio.load_case(os.path.join(os.path.dirname(EIN91), name), 100)
    ctx = Context(c.tables)
    dc = DeviceCase(ctx, c)
    dc.upload(c.P)
    chosen = dc.tune(path, tile)
    rep = dc.step(100)
    Pg = dc.download()
    dc.destroy()
    ctx.close()
    Po = c.P.copy()
    dts = OracleCase(c).step(Po, 100)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    a, b = c.layout.interior(Pg, 0), c.layout.interior(Po, 0)
    err = max(np.abs(a[s] - b[s]).max() / max(np.abs(b[s]).max(), 1.0) for s in range(c.nprim))
    print(f"{name} HLLE (reference inputs), path {path} tile {chosen} x 100: rel Linf {err:.3e}, values differing "
          f"bitwise {int(np.sum(a != b))} of {a.size}", flush=True)
    assert err <= TOL


def test_run_to_time_threshold(gpu_ready):
    """caseio.run_case on the reference's test/1D/Noh87 inputs, which end on time-threshold (no iter-threshold): the
    device steps while its time is below the threshold, as the reference's loop does, and lands on the oracle's step
    count, time and state."""
    c = caseio.load_case(os.path.join(os.path.dirname(EIN91), "noh87"))
    t_end = c.meta["time_threshold"]
    assert c.nsteps == 0 and t_end > 0
    ctx = Context(c.tables)
    dc = DeviceCase(ctx, c)
    dc.upload(c.P)
    rep = caseio.run_case(dc, c)
    Pg = dc.download()
    dc.destroy()
    ctx.close()
    o = OracleCase(c)
    Po = c.P.copy()
    t, n = 0.0, 0
    while t < t_end:
        t += o.step(Po, 1)[0]
        n += 1
    assert rep.steps == n and rep.time >= t_end and abs(rep.time - t) <= 1e-13 * t
    a, b = c.layout.interior(Pg, 0), c.layout.interior(Po, 0)
    err = max(np.abs(a[s] - b[s]).max() / max(np.abs(b[s]).max(), 1.0) for s in range(c.nprim))
    print(f"noh87 to time-threshold {t_end}: {n} steps, time {rep.time:.6f}, rel Linf {err:.3e}", flush=True)
    assert err <= TOL
