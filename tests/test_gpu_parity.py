"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same inputs.

Bar: relative L-infinity <= 1e-12 on the conservative variables after N steps (the tolerance of the residual path);
both sides are built without FMA contraction and keep the reference's rounding order, so the results are expected
to be bitwise equal, which the tests also report."""
import numpy as np
import pytest

from mosegpu import Context, DeviceCase, cases
from oracle_bind import OracleCase

pytestmark = pytest.mark.gpu

TOL = 1e-12


def conservative(case, P):
    out = []
    for b in range(case.layout.nb):
        V = case.layout.interior(P, b).reshape(case.nprim, -1)
        ns = case.scheme.nsc
        rho = V[:ns].sum(axis=0)
        u, v, w, p = V[ns], V[ns + 1], V[ns + 2], V[ns + 3]
        # E with the table h = cp T of the air cases: e = cv T = p / (rho (gamma - 1)) up to the table
        E = p / (case.meta.get("gamma", 1.4) - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
        out.append(np.vstack([V[:ns], rho * u, rho * v, rho * w, E, V[ns + 4:]]))
    return np.hstack(out)


def rel_linf(a, b):
    d = np.abs(a - b).max(axis=1)
    s = np.abs(b).max(axis=1)
    s[s == 0] = 1.0
    return (d / s).max()


def run_pair(case, nsteps):
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    rep = dc.step(nsteps)
    Pg = dc.download()
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, nsteps)
    return Pg, Po, rep, dts, ctx


def test_sod_parity(gpu_ready):
    case = cases.sod(1000, 200)
    case.meta["gamma"] = 1004.5 / (1004.5 - 8314.46261815324 / 28.970418)
    Pg, Po, rep, dts, _ = run_pair(case, 200)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    err = rel_linf(conservative(case, Pg), conservative(case, Po))
    nbit = int(np.sum(case.layout.interior(Pg, 0) != case.layout.interior(Po, 0)))
    print(f"sod 1000 x 200 steps: rel Linf {err:.3e}, cells differing bitwise {nbit}")
    assert err <= TOL


def test_sod_residual_and_dt(gpu_ready):
    case = cases.sod(200, 0)
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    Rg = dc.residual()
    Po = case.P.copy()
    o = OracleCase(case)
    Ro = o.residual(Po)
    Ig, Io = case.layout.interior(Rg, 0), case.layout.interior(Ro, 0)
    assert np.abs(Ig - Io).max() <= TOL * np.abs(Io).max()
    assert dc.dt() == o.dt(case.P.copy())


@pytest.mark.parametrize("n,steps", [(16, 10), (24, 4)])
def test_taylor_green_parity(gpu_ready, n, steps):
    case = cases.taylor_green(n, steps)
    case.meta["gamma"] = 1004.5 / (1004.5 - 8314.46261815324 / 28.970418)
    Pg, Po, rep, dts, _ = run_pair(case, steps)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    err = rel_linf(conservative(case, Pg), conservative(case, Po))
    nbit = int(np.sum(case.layout.interior(Pg, 0) != case.layout.interior(Po, 0)))
    print(f"taylor-green {n}^3 x {steps} steps: rel Linf {err:.3e}, cells differing bitwise {nbit}")
    assert err <= TOL


def test_taylor_green_residual(gpu_ready):
    case = cases.taylor_green(20, 0, dims=(20, 12, 16))
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    Rg = dc.residual()
    Ro = OracleCase(case).residual(case.P.copy())
    Ig, Io = case.layout.interior(Rg, 0), case.layout.interior(Ro, 0)
    for s in range(5):
        assert np.abs(Ig[s] - Io[s]).max() <= TOL * np.abs(Io[s]).max(), s


def test_gpu_errors_are_reported(gpu_ready):
    case = cases.sod(64, 1)
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    P = case.P.copy()
    case.layout.interior(P, 0)[4, 0, 0, 40] = -5.0   # negative pressure in one cell: the update fails the state check
    dc.upload(P)
    with pytest.raises(Exception) as ei:
        dc.step(1)
    assert "MOSE_ERR_STATE" in str(ei.value)


# --------------------------------------------------------------------------------------------------------------
# the fused stage kernel (default path) against the oracle, on tile shapes that do and do not divide the block
# --------------------------------------------------------------------------------------------------------------
def run_fused(case, nsteps, tile):
    import mosegpu
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    chosen = dc.tune(mosegpu.PATH_FUSED, tile)
    rep = dc.step(nsteps)
    return dc.download(), rep, chosen


@pytest.mark.parametrize("tile", [(32, 4, 2), (16, 4, 4), (64, 2, 2), (0, 0, 0), (8, 8, 8), (32, 4, 4), (8, 4, 4)])
def test_fused_taylor_green_parity(gpu_ready, tile):
    case = cases.taylor_green(20, 6, dims=(20, 18, 21))
    case.meta["gamma"] = 1004.5 / (1004.5 - 8314.46261815324 / 28.970418)
    Pg, rep, chosen = run_fused(case, 6, tile)
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, 6)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    err = rel_linf(conservative(case, Pg), conservative(case, Po))
    nbit = int(np.sum(case.layout.interior(Pg, 0) != case.layout.interior(Po, 0)))
    print(f"fused tile {chosen}: taylor-green 20x18x21 x 6 steps rel Linf {err:.3e}, bitwise-differing {nbit}")
    assert err <= TOL


def test_fused_sod_parity(gpu_ready):
    case = cases.sod(1000, 200)
    case.meta["gamma"] = 1004.5 / (1004.5 - 8314.46261815324 / 28.970418)
    Pg, rep, chosen = run_fused(case, 200, (256, 1, 1))
    Po = case.P.copy()
    OracleCase(case).step(Po, 200)
    err = rel_linf(conservative(case, Pg), conservative(case, Po))
    print(f"fused sod 1000 x 200: rel Linf {err:.3e}")
    assert err <= TOL


def test_fused_equals_plain_at_bench_size(gpu_ready):
    """At the benchmark's full size (256^3) the oracle is too slow: the fused kernel must give the bits of the
    per-face kernels, and the total mass on the periodic box must stay constant (to rounding)."""
    import mosegpu
    case = cases.taylor_green(256, 2)
    ctx = Context(case.tables)
    out = {}
    for name, path, tile in (("plain", mosegpu.PATH_PLAIN, (0, 0, 0)), ("fused", mosegpu.PATH_FUSED, (32, 4, 2)),
                             ("fused512", mosegpu.PATH_FUSED, (8, 8, 8))):
        dc = DeviceCase(ctx, case)
        dc.upload(case.P)
        dc.tune(path, tile)
        dc.step(2)
        out[name] = dc.download()
        dc.destroy()
    a = case.layout.interior(out["plain"], 0)
    b = case.layout.interior(out["fused"], 0)
    assert np.array_equal(a, b)
    assert np.array_equal(a, case.layout.interior(out["fused512"], 0))
    vol = case.layout.interior(case.geom.vol, 0)
    m0 = np.sum(case.layout.interior(case.P, 0)[0] * vol)
    m1 = np.sum(b[0] * vol)
    assert abs(m1 - m0) <= 1e-12 * m0


# --------------------------------------------------------------------------------------------------------------
# residual norm (Compute_Residual_blk) and state hash (field_state_hash)
# --------------------------------------------------------------------------------------------------------------
def oracle_residual_norm(case, P_new, P_old):
    """sqrt of sum over the interior cells of (species-sum change)^2, (u, v, w, p, RANS change)^2."""
    ns = case.scheme.nsc
    tot = np.zeros(case.nprim - ns + 1)
    for b in range(case.layout.nb):
        a = case.layout.interior(P_new, b).reshape(case.nprim, -1)
        o = case.layout.interior(P_old, b).reshape(case.nprim, -1)
        d = (a - o)
        tot[0] += np.sum(d[:ns].sum(axis=0) ** 2)
        tot[1:] += np.sum(d[ns:] ** 2, axis=1)
    return np.sqrt(tot)


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_residual_norm_and_hash(gpu_ready, path):
    import mosegpu
    case = cases.taylor_green(16, 2)
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    assert dc.state_hash() == mosegpu.state_hash(case.layout, case.P)
    dc.tune(mosegpu.PATH_FUSED if path == "fused" else mosegpu.PATH_PLAIN, (0, 0, 0))
    dc.step(1)
    P1 = dc.download()
    dc.step(1)
    P2 = dc.download()
    rn = dc.residual_norm()
    ref = oracle_residual_norm(case, P2, P1)
    assert np.allclose(rn, ref, rtol=1e-12, atol=0), (rn, ref)
    assert dc.state_hash() == mosegpu.state_hash(case.layout, P2)
    # the oracle's state after two steps is bitwise the device's: same hash
    Po = case.P.copy()
    OracleCase(case).step(Po, 2)
    assert mosegpu.state_hash(case.layout, Po) == dc.state_hash()


def test_graph_replays_equal_eager_steps(gpu_ready):
    """The captured fused step, replayed step after step with host work and downloads in between, gives the bits of
    the eagerly launched steps."""
    import mosegpu
    case = cases.taylor_green(20, 8, dims=(20, 16, 12))
    ctx = Context(case.tables)
    runs = {}
    for name, graph in (("graph", True), ("eager", False)):
        dc = DeviceCase(ctx, case)
        dc.upload(case.P)
        dc.tune(mosegpu.PATH_FUSED, (16, 4, 4), graph=graph)
        states = []
        for _ in range(4):
            dc.step(2)
            states.append(dc.download())
            _ = np.sum(states[-1])          # host work between replays
        rep = dc.step(3)
        states.append(dc.download())
        runs[name] = (states, rep.time, dc.residual_norm())
        dc.destroy()
    for a, b in zip(runs["graph"][0], runs["eager"][0]):
        assert np.array_equal(case.layout.interior(a, 0), case.layout.interior(b, 0))
    assert runs["graph"][1] == runs["eager"][1]
    assert np.array_equal(runs["graph"][2], runs["eager"][2])
