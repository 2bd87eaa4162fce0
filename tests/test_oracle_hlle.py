"""Pins of the oracle's HLLE flux (rs 2: conv_physics.inc:116-124, Lib_Riemann_HLL.f90:137-198) on the CPU.

- Consistency: equal states give the physical flux (rho un, rho u un + p n, rho H un).
- Upwinding: when every signal speed is positive (S1 = 0) the flux is the left state's physical flux, and when every
  one is negative (S4 = 0) the right state's.
- The reference's own test/1D/Ein91 inputs (Einfeldt's 1-2-0-3 double rarefaction, HLLE + van Leer, RK3 at CFL 0.9)
  run to t = 0.15 within an L1 density error of 0.02 of the exact solution.
- The reference's own test/1D/Noh87 inputs (HLLE, two cold streams colliding): the shocked gas holds the closed-form
  density (gamma + 1) / (gamma - 1) = 6 between the two shocks, which run outward at (gamma - 1) / 2 |u0|.
"""
import os

import numpy as np
import pytest

from exact_riemann import sample
from mosegpu import caseio, cases
from oracle_bind import OracleCase

EIN91 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "ein91")
GAM = cases.CP_AIR / (cases.CP_AIR - cases.R_AIR)


def physical(q, n):
    rho, u, p = q[0], np.array(q[1:4]), q[4]
    un = u @ n
    H = GAM / (GAM - 1.0) * p / rho + 0.5 * u @ u
    return np.array([rho * un, *(rho * u * un + p * n), rho * H * un])


@pytest.fixture(scope="module")
def oracle():
    c = cases.box_case((4, 4, 4), {f: "periodic" for f in range(1, 7)}, viscous=False, riemann="HLLE")
    return OracleCase(c)


def test_consistency_and_upwinding(oracle):
    rng = np.random.Generator(np.random.PCG64(8))
    for _ in range(50):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        q = np.array([rng.uniform(0.2, 3.0), *rng.uniform(-200, 200, 3), rng.uniform(2e4, 3e5)])
        F = oracle.riemann(2, q, q, n)
        assert np.allclose(F, physical(q, n), rtol=1e-12, atol=1e-9 * np.abs(physical(q, n)).max())
    a = np.sqrt(GAM * 1e5 / 1.0)
    n = np.array([1.0, 0.0, 0.0])
    qL = np.array([1.0, 3.0 * a, 10.0, 0.0, 1e5])
    qR = np.array([0.8, 2.8 * a, -5.0, 1.0, 0.9e5])
    assert np.allclose(oracle.riemann(2, qL, qR, n), physical(qL, n), rtol=1e-13)
    qL[1], qR[1] = -3.0 * a, -2.9 * a
    assert np.allclose(oracle.riemann(2, qL, qR, n), physical(qR, n), rtol=1e-13)


def _ein91_error(c, nsteps_max=4000):
    o = OracleCase(c)
    P = c.P.copy()
    t = 0.0
    for _ in range(nsteps_max):
        if t >= 0.15 - 1e-12:
            break
        t += o.step(P, 1)[0]
    x = 0.5 * (c.nodes[0][0, 0, 1:, 0] + c.nodes[0][0, 0, :-1, 0])
    V0 = c.layout.interior(c.P, 0)[:, 0, 0, :]
    L, R = V0[:, 0], V0[:, -1]
    x0 = x[np.argmax(np.diff(V0[1]) != 0) + 1] - 0.5 * (x[1] - x[0])
    rho, u, p = sample(x, t, x0, L[0], L[1], L[4], R[0], R[1], R[4], g=GAM)
    W = c.layout.interior(P, 0)[:, 0, 0, :]
    return t, np.mean(np.abs(W[0] - rho)) , (x[-1] - x[0])


def test_ein91_against_exact():
    c = caseio.load_case(EIN91, 1)
    s = c.scheme
    assert (s.riemann, s.limiter, s.n_rk, s.cfl) == (2, 2, 3, 0.9)
    t, err, _ = _ein91_error(c)
    assert abs(t - 0.15) < 0.01
    assert err < 0.02, err


NOH87 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "noh87")


def test_noh87_shocked_density():
    c = caseio.load_case(NOH87, 200)
    P = c.P.copy()
    t = OracleCase(c).step(P, 200).sum()
    x = 0.5 * (c.nodes[0][0, 0, 1:, 0] + c.nodes[0][0, 0, :-1, 0])
    V0 = c.layout.interior(c.P, 0)[:, 0, 0, :]
    u0 = abs(V0[1, 0])
    x0 = x[np.argmax(np.diff(V0[1]) != 0) + 1] - 0.5 * (x[1] - x[0])
    W = c.layout.interior(P, 0)[:, 0, 0, :]
    xs = (GAM - 1.0) / 2.0 * u0 * t                    # shock distance from the centre
    band = (np.abs(x - x0) > 0.3 * xs) & (np.abs(x - x0) < 0.8 * xs)
    assert band.sum() >= 6
    assert abs(np.median(W[0][band]) - (GAM + 1.0) / (GAM - 1.0)) <= 0.03 * 6.0
    assert np.all(np.abs(W[0][np.abs(x - x0) > 1.3 * xs + 0.02] - 1.0) < 0.05)   # unshocked gas ahead of the shocks
