"""The performance path (MOSE_PATH_FAST: the fused stage with contracted multiply-adds, reciprocals and the short
form of the Roe flux) against the oracle on every family of parity cases of the fused path.

The parity path equals the oracle bit for bit; this path differs from it by rounding only, and the bar is the
residual path's: a relative L-infinity <= 1e-12 per primitive variable after N steps (<= 1e-8 with chemistry, whose
adaptive step counts turn a rounding into a different step sequence), the velocity in units of |u| + sqrt(p / rho)
(fast_error).  Noh87 is held to 1e-9: a change of rounding alone moves it by ~2e-11 (test_oracle_rounding.py).
Each comparison reports how many values differ bitwise; the time, the sum of the steps' dt, is held to the same bar.
"""
import os

import numpy as np
import pytest

import mosegpu
from mosegpu import caseio, cases, chem
from oracle_bind import OracleCase, OracleChem
from test_gpu_walls_sa import run_device
from test_oracle_species import DIFF1, Y4, as_four_species
from test_oracle_symmetry import mirror_pair

pytestmark = pytest.mark.gpu

TOL = 1e-12
FAST = mosegpu.PATH_FAST
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def fast_error(case, Pg, Po, what):
    """Relative L-infinity per primitive variable over every block's interior.  The three velocity components share
    one scale, the largest |u| + sqrt(p / rho) anywhere (sqrt(p / rho) <= a for every gamma >= 1): a rounding of the
    pressure moves the velocity by about (dp / p) a (the acoustic relation du = dp / (rho a)), so the velocity's
    rounding is measured in that unit, also where the flow is at rest or slow (v and w of a 1D case are round-off;
    diff1's diffusion-driven velocities are far below a).  The printout also gives the error against max |u|."""
    ns = case.scheme.nsc
    a = np.hstack([case.layout.interior(Pg, b).reshape(case.nprim, -1) for b in range(case.layout.nb)])
    b = np.hstack([case.layout.interior(Po, b).reshape(case.nprim, -1) for b in range(case.layout.nb)])
    scale = np.abs(b).max(axis=1)
    rho = b[:ns].sum(axis=0)
    speed = np.sqrt((b[ns:ns + 3] ** 2).sum(axis=0))
    snd = np.sqrt(b[ns + 3] / rho)
    u_max = float(np.abs(b[ns:ns + 3]).max())
    scale[ns:ns + 3] = float((speed + snd).max())
    scale[scale == 0] = 1.0
    err = np.abs(a - b).max(axis=1) / scale
    du = float(np.abs(a[ns:ns + 3] - b[ns:ns + 3]).max())
    worst = int(np.argmax(err))
    print(f"{what}: rel Linf {err.max():.3e} (variable {worst}; velocity against max |u| {du / max(u_max, 1e-300):.3e}),"
          f" values differing bitwise {int(np.sum(a != b))} of {a.size}", flush=True)
    return float(err.max())


def check(case, nsteps, tile=(0, 0, 0), what="", tol=TOL):
    Pg, rep, chosen = run_device(case, nsteps, FAST, tile)
    Po = case.P.copy()
    dts = OracleCase(case, nthreads=16).step(Po, nsteps)
    err = fast_error(case, Pg, Po, f"fast path: {what} tile {chosen} x {nsteps}")
    dtime = abs(rep.time - dts.sum()) / dts.sum()
    print(f"  time (the sum of the steps' dt): relative difference {dtime:.3e}", flush=True)
    assert err <= tol
    assert dtime <= tol


def test_fast_sod(gpu_ready):
    check(cases.sod(1000, 200), 200, what="sod 1000 (HLLC, minmod)")


@pytest.mark.parametrize("tile", [(0, 0, 0), (8, 8, 4), (8, 8, 8), (16, 4, 4)])
def test_fast_taylor_green(gpu_ready, tile):
    check(cases.taylor_green(24, 10), 10, tile, "taylor-green 24^3 (LMRoe, van Albada, viscous)")


def test_fast_taylor_green_odd_box(gpu_ready):
    check(cases.taylor_green(20, 6, dims=(20, 18, 21)), 6, (8, 8, 4), "taylor-green 20x18x21")


@pytest.mark.parametrize("rs", ["HLLC", "AUSM+", "LMRoe", "HLLE"])
def test_fast_channel_lattice(gpu_ready, rs):
    case = cases.channel_lattice(nb=(2, 2, 2), wmin=4, wmax=9, seed=11, nsteps=5, riemann=rs)
    check(case, 5, (0, 0, 0), f"channel 2x2x2 (walls, far field, SA, {rs})")


def test_fast_channel_lattice_4x4x4(gpu_ready):
    case = cases.channel_lattice(nb=(4, 4, 4), wmin=6, wmax=12, seed=5, nsteps=3, riemann="LMRoe")
    check(case, 3, (8, 4, 4), "channel 4x4x4 (64 blocks, LMRoe)")


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("viscous", [False, True])
def test_fast_symmetry(gpu_ready, axis, viscous):
    case, _ = mirror_pair(axis, [12, 10, 9], viscous, riemann="HLLC" if viscous else "LMRoe", seed=31 + axis)
    check(case, 4, (0, 0, 0), f"symmetry axis {axis} viscous {viscous}")


def test_fast_four_species(gpu_ready):
    c1 = cases.box_case((12, 10, 9), {f: "periodic" for f in range(1, 7)}, extent=(1.0, 0.9, 0.8), viscous=True,
                        mu_of_T=cases.sutherland(1.8e-3, 300.0), riemann="LMRoe", limiter="vanleer")
    rng = np.random.Generator(np.random.PCG64(13))
    V = c1.layout.interior(c1.P, 0)
    V[0] = 1.1 + 0.1 * rng.uniform(-1, 1, V[0].shape)
    for q in range(3):
        V[1 + q] = 30.0 * rng.uniform(-1, 1, V[0].shape)
    V[4] = 1e5 * (1.0 + 0.05 * rng.uniform(-1, 1, V[0].shape))
    cases.fill_ghosts_nearest(c1.layout, c1.P)
    Yr = Y4 * (1.0 + 0.2 * rng.uniform(-1, 1, 4))
    case = as_four_species(c1, Yr / Yr.sum(), "LMRoe")
    W = case.layout.interior(case.P, 0)
    W[0] *= 1.0 + 0.3 * rng.uniform(-1, 1, W[0].shape)
    cases.fill_ghosts_nearest(case.layout, case.P)
    check(case, 4, (0, 0, 0), "four-species viscous box (LMRoe)")


def test_fast_reference_cases(gpu_ready):
    """The reference's own inputs: diff1 (four species, symmetry), the isentropic vortex, Shu-Osher, Ein91, Noh87."""
    check(caseio.load_case(DIFF1, 100), 100, (0, 0, 0), "diff1 (reference inputs)")
    check(caseio.load_case(os.path.join(GOLD, "isentropic-vortex"), 20), 20, (0, 0, 0), "isentropic vortex")
    check(caseio.load_case(os.path.join(GOLD, "shu-osher"), 330), 330, (0, 0, 0), "Shu-Osher (reference inputs)")
    check(caseio.load_case(os.path.join(GOLD, "ein91"), 100), 100, (0, 0, 0), "ein91 (reference inputs, HLLE)")
    # Noh87 moves by ~2e-11 (time ~2e-12) under a change of rounding alone: the oracle compiled with contraction
    # against the oracle (test_oracle_rounding.py); its bar is 1e-9
    check(caseio.load_case(os.path.join(GOLD, "noh87"), 100), 100, (0, 0, 0), "noh87 (reference inputs, HLLE)", 1e-9)


def test_fast_riemann2d(gpu_ready):
    check(cases.riemann2d(64, 40, seed=3), 40, (0, 0, 0), "2D four-quadrant Riemann 64^2")


@pytest.mark.skip(reason="a GPU memory fault (aperture violation, reported in k_bcrec<9,0>) during this test's autotune "
                         "on a build with a batched phase-L load loop for nine species and a compacted record list; no "
                         "cause was found by reading the code, both changes were taken out (the nine-species kernels are "
                         "again instruction for instruction those of the build that passed this test), and the case "
                         "stays off the GPU until a cause is shown")
def test_fast_reacting_box(gpu_ready):
    from mosegpu import Context, DeviceCase, DeviceChem
    mech = chem.h2_air_mechanism()
    case = cases.reacting_box((12, 10, 8), viscous=True, T_range=(1200.0, 1700.0))
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dch = DeviceChem(ctx, mech, 1)
    dc.upload(case.P)
    chosen = dc.tune(FAST, (0, 0, 0))
    # the integrator's tolerance sets how far a rounding of the transport moves the chemistry's adaptive steps: at
    # rtol = 1e-8 a step decision that flips is a change of the bar's own size, so the check runs at rtol = 1e-10
    dc.chem_attach(dch, chem.ode_opts(rtol=1e-10, atol=1e-14))
    rep = dc.step(3)
    Pg = dc.download()
    dc.destroy(); dch.destroy(); ctx.close()
    o = OracleCase(case)
    o.set_chemistry(OracleChem(mech, nthreads=16), rtol=1e-10, atol=1e-14)
    Po = case.P.copy()
    dts = o.step(Po, 3)
    # the chemistry's adaptive steps turn the transport's roundings into differences at the integration tolerance,
    # and the next step's dt follows the state
    assert abs(rep.time - dts.sum()) <= 1e-8 * dts.sum()
    assert fast_error(case, Pg, Po, f"fast path: reacting box 12x10x8 tile {chosen} x 3 Strang steps") <= 1e-8
