"""Pins of the oracle's Strang-split reacting step (Explicit_Step's strang loop, Mod_Explicit.f90:155-294, with
set_StrangSplitting, Lib_Strang.f90:10-22, and Chemistry_Newstate_blk, Lib_Chemistry.f90:33-76) on the CPU.

- With a mechanism of no reactions the chemistry leaves every cell's species and temperature as they are, so one
  Strang step is two SSP-RK3 transport substeps of strangcoeff 1/2 with the step's one dt (the second substage does
  not recompute it), each from the state the previous one left, then p = T rho R again at the end of the chemistry:
  the step must equal that sequence written out with the oracle's own residual and update on dt/2 (multiplying by
  1/2 is exact, so the bits agree), and differ from one full step.
- With the hydrogen-air mechanism the step conserves each element's mass on the periodic box to 1e-11 (the rates
  conserve elements reaction by reaction and a Rosenbrock step keeps linear invariants; transport conserves every
  species), and the chemistry is the batch chemistry of the same cells (gathered from the arena's species rows and
  p row).
"""
import numpy as np

from mosegpu import cases, chem
from oracle_bind import OracleCase, OracleChem

MECH = chem.h2_air_mechanism()
ELEMENTS = {"H": [2, 0, 2, 1, 0, 1, 1, 2, 0], "O": [0, 2, 1, 0, 1, 1, 2, 2, 0], "N": [0, 0, 0, 0, 0, 0, 0, 0, 2]}


def no_reactions():
    m = chem.h2_air_mechanism()
    m.ni1 = np.zeros((0, m.ns + 1)); m.ni2 = np.zeros((0, m.ns + 1)); m.eps = np.zeros((0, m.ns + 1))
    m.kf = np.zeros((0, chem.NTAB)); m.kb = np.zeros((0, chem.NTAB))
    m.troe = np.zeros((0, 4, chem.NTAB)); m.lind = np.zeros((0, 3, chem.NTAB))
    return m


def p_recomputed(c, P):
    """p = T rho R with T = p / (rho R) (Chemistry_Newstate_blk's entry and exit on an unchanged cell), in the
    oracle's order of operations."""
    Ri = chem.RU / chem.WM
    V = c.layout.interior(P, 0)
    rho = V[0].copy()
    for s in range(1, 9):
        rho = rho + V[s]
    R = np.zeros_like(rho)
    for s in range(9):
        R = R + Ri[s] * V[s] / rho
    T = V[12] / (rho * R)
    V[12] = T * rho * R


def test_strang_step_with_no_reactions_is_two_half_steps():
    c = cases.reacting_box((8, 6, 5))
    o = OracleCase(c)
    oc = OracleChem(no_reactions(), nthreads=4)
    P = c.P.copy()
    dt = o.dt(P)
    # the written-out sequence: two substeps of SSP-RK3 on dt/2, p recomputed after the first
    rk = (1.0, 0.25, 2.0 / 3.0)
    Q = c.P.copy()
    for sub in range(2):
        Q0 = Q.copy()
        for irk in range(3):
            R = o.residual(Q, dt)
            o.newstate(Q, Q0, R, dt / 2, rk[irk])
        if sub == 0:
            p_recomputed(c, Q)
    o.set_chemistry(oc)
    got = o.step(P, 1)
    assert got[0] == dt
    Vi, Wi = c.layout.interior(P, 0), c.layout.interior(Q, 0)
    assert np.array_equal(Vi, Wi), np.max(np.abs(Vi - Wi))
    # and not one full transport step
    o.set_chemistry(None)
    F = c.P.copy()
    o.step(F, 1)
    assert np.max(np.abs(c.layout.interior(F, 0)[9] - Vi[9])) > 1e-6 * np.max(np.abs(Vi[9]))


def element_mass(c, P):
    V = c.layout.interior(P, 0)
    vol = c.layout.interior(c.geom.vol.reshape(1, -1), 0)[0]
    return {e: sum(a * np.sum(V[s] * vol) / chem.WM[s] for s, a in enumerate(w)) for e, w in ELEMENTS.items()}


def test_strang_step_conserves_elements():
    c = cases.reacting_box((8, 6, 5), T_range=(1200.0, 1700.0))
    o = OracleCase(c)
    o.set_chemistry(OracleChem(MECH, nthreads=8))
    P = c.P.copy()
    m0 = element_mass(c, P)
    o.step(P, 3)
    assert o.chem_failed() == 0
    m1 = element_mass(c, P)
    for e in m0:
        assert abs(m1[e] - m0[e]) <= 1e-11 * m0[e], (e, (m1[e] - m0[e]) / m0[e])
    # the chemistry did something: water formed
    V0, V1 = c.layout.interior(c.P, 0), c.layout.interior(P, 0)
    assert np.sum(V1[2]) > np.sum(V0[2]) * 1.5 + 1e-12


def test_case_chemistry_is_the_batch_chemistry():
    c = cases.reacting_box((6, 5, 4), T_range=(1300.0, 1800.0))
    o = OracleCase(c)
    oc = OracleChem(MECH, nthreads=8)
    o.set_chemistry(oc)
    P = c.P.copy()
    o.react(P, 2e-6)
    V = c.layout.interior(c.P, 0).reshape(13, -1)
    st = np.ascontiguousarray(np.concatenate([V[:9], V[12:13]]))
    ref, _, idid, _ = oc.advance(st, 2e-6)
    assert np.all(idid == 1)
    W = c.layout.interior(P, 0).reshape(13, -1)
    assert np.array_equal(W[:9], ref[:9]) and np.array_equal(W[12], ref[9])
    assert np.array_equal(W[9:12], V[9:12])          # velocities untouched
