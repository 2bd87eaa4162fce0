"""Parity of the device chemistry (persistent RODAS and RADAU5 kernels through the C ABI) with the chemistry oracle.

The north star's bar for chemistry is a relative L-infinity <= 1e-8 per variable.  The device's controller powers
are glibc's pow bit for bit (mose_pow: checked here on the device against the host's libm), and the oracle equals
the reference's compiled Hairer integrators bit for bit (test_oracle_vs_hairer.py), so the device must equal the
oracle exactly: same states, same idid, same step counts.  At the larger batch size the checks are size-free
invariants: the species mass.
"""
import numpy as np
import pytest

from mosegpu import Context, DeviceChem, chem
from oracle_bind import OracleChem

pytestmark = pytest.mark.gpu

TOL = 1e-8
MECH = chem.h2_air_mechanism()


def run_device(state, dt, rtol=1e-8, atol=1e-12, method=1):
    ctx = Context(chem_tables())
    dc = DeviceChem(ctx, MECH, state.shape[1])
    dc.upload(state)
    rep = dc.advance(dt, chem.ode_opts(rtol, atol, method=method))
    out = dc.download()
    dc.destroy()
    ctx.close()
    return out, rep


def chem_tables():
    from mosegpu.core import ThermoTables
    return ThermoTables(MECH.wm, MECH.cp, MECH.h, names=MECH.names)


def test_chem_parity_with_oracle(gpu_ready):
    st = chem.h2_air_cells(384, seed=21)
    (dev, hr, idid, nstep), rep = run_device(st, 1e-6)
    ref, hr_o, idid_o, nstep_o = OracleChem(MECH, nthreads=16).advance(st, 1e-6)
    assert np.all(idid == 1) and np.all(idid_o == 1)
    scale = np.abs(ref).max(axis=1)
    err = (np.abs(dev - ref).max(axis=1) / scale).max()
    same_steps = float(np.mean(nstep == nstep_o))
    print(f"chemistry 384 cells: rel Linf {err:.3e}, identical step counts {same_steps:.3f}, "
          f"steps {rep.steps_total}, kernel {rep.ms:.1f} ms", flush=True)
    assert err <= TOL
    assert np.array_equal(dev, ref) and np.array_equal(nstep, nstep_o)
    assert np.array_equal(hr, hr_o)


def test_chem_invariants_large_batch(gpu_ready):
    st = chem.h2_air_cells(65536, seed=22)
    (dev, hr, idid, nstep), rep = run_device(st, 1e-6)
    assert np.all(idid == 1)
    m0, m1 = st[:9].sum(0), dev[:9].sum(0)
    assert np.max(np.abs(m1 - m0) / m0) <= 1e-10
    print(f"chemistry 65536 cells: {rep.ms:.1f} ms, {65536 / (rep.ms * 1e-3):.3e} cells/s, steps mean "
          f"{nstep.mean():.1f} max {nstep.max()}, rhs {rep.rhs_total}", flush=True)


def test_radau5_parity_with_oracle(gpu_ready):
    """RADAU5 (method 2) through the same C ABI: the cells where the simplified Newton iteration stalls and the
    integration stops at NMAX (idid -2, as in the reference's own RADAU5) must stop on the device too."""
    st = chem.h2_air_cells(384, seed=21)
    (dev, hr, idid, nstep), rep = run_device_allow_fail(st, 1e-6, method=2)
    ref, hr_o, idid_o, nstep_o = OracleChem(MECH, nthreads=16).advance(st, 1e-6, method=2)
    assert np.array_equal(idid, idid_o)
    ok = idid == 1
    assert ok.mean() > 0.8
    first = int(np.argmax(~ok))
    assert rep.failed == int((~ok).sum()) and rep.first_failed_cell == first and rep.first_failed_idid == idid[first]
    scale = np.abs(ref[:, ok]).max(axis=1)
    err = (np.abs(dev[:, ok] - ref[:, ok]).max(axis=1) / scale).max()
    same_steps = float(np.mean(nstep == nstep_o))
    print(f"radau5 384 cells: rel Linf {err:.3e}, identical step counts {same_steps:.3f}, failed {int((~ok).sum())}, "
          f"kernel {rep.ms:.1f} ms", flush=True)
    assert err <= TOL
    assert np.array_equal(dev, ref) and np.array_equal(nstep, nstep_o)
    assert np.array_equal(hr[ok], hr_o[ok])


def run_device_allow_fail(state, dt, rtol=1e-8, atol=1e-12, method=2):
    """As run_device, for a batch in which some cells may end with idid < 0 (MOSE_ERR_ODE is then expected)."""
    from mosegpu.core import MoseError
    ctx = Context(chem_tables())
    dc = DeviceChem(ctx, MECH, state.shape[1])
    dc.upload(state)
    try:
        rep = dc.advance(dt, chem.ode_opts(rtol, atol, method=method))
    except MoseError as e:
        assert "cells failed" in str(e)
        rep = dc.last_report
    out = dc.download()
    dc.destroy()
    ctx.close()
    return out, rep


@pytest.mark.parametrize("n", [1, 13, 61])
def test_chem_ragged_batches(gpu_ready, n):
    """Batches that fill neither a team row nor a workgroup (1, 13 and 61 cells; 6 teams per wavefront, 24 per
    workgroup): the oracle's states bit for bit."""
    st = chem.h2_air_cells(n, seed=100 + n)
    (dev, hr, idid, nstep), rep = run_device(st, 1e-6)
    ref, hr_o, idid_o, nstep_o = OracleChem(MECH, nthreads=4).advance(st, 1e-6)
    assert rep.cells == n
    assert np.array_equal(idid, idid_o) and np.array_equal(nstep, nstep_o)
    assert np.array_equal(dev, ref)


def test_chem_argument_errors(gpu_ready):
    from mosegpu.core import MoseError
    ctx = Context(chem_tables())
    dc = DeviceChem(ctx, MECH, 8)
    with pytest.raises(MoseError):
        dc.advance(1e-6, chem.ode_opts())          # nothing uploaded
    with pytest.raises(MoseError):
        dc.upload(chem.h2_air_cells(9, seed=1))    # more cells than the object holds
    dc.upload(chem.h2_air_cells(8, seed=1))
    with pytest.raises(MoseError):
        dc.advance(1e-6, chem.ode_opts(method=7))  # no such integrator
    with pytest.raises(MoseError):
        dc.advance(-1e-6, chem.ode_opts())         # dt must be positive
    dc.destroy()
    ctx.close()


@pytest.mark.parametrize("kind", ["controllers", "troe", "general", "bits", "negative_integer_y", "subnormal_x",
                                  "near_over_underflow", "near_one"])
def test_device_pow_is_the_host_pow(gpu_ready, kind):
    """The device's pow (mose_pow, compiled for gfx950) against the host's libm pow on 1e6 arguments of each class:
    the same bits."""
    from oracle_bind import _d, oracle_lib
    from test_libm import KINDS, arguments, same_bits
    rng = np.random.default_rng(1000 + KINDS.index(kind))
    x, y = arguments(kind, 1_000_000, rng)
    ctx = Context(chem_tables())
    dev = ctx.pow_batch(x, y)
    ctx.close()
    host = np.empty_like(x)
    oracle_lib().oc_pow_batch(1, x.size, _d(x), _d(y), _d(host))
    bad = ~same_bits(dev, host)
    print(f"device pow, {kind}: {int(bad.sum())} of {x.size} differ from the host libm", flush=True)
    assert bad.sum() == 0, (x[bad][:4], y[bad][:4], dev[bad][:4], host[bad][:4])
