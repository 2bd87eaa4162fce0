"""Pins of the oracle's whole step (CPU): the Sod problem against Toro's exact solution (first-order-in-dx
convergence of the L1 error, the star-region plateau), conservation of the totals while the waves are inside
the tube, the exact zero residual of a uniform state, discrete conservation on the periodic box, and the
SSP-RK3 amplification factor on a linear problem."""
import numpy as np
import pytest

from exact_riemann import sample, star_state
from mosegpu import cases
from oracle_bind import OracleCase


def run_sod(n, t_end):
    c = cases.sod(n, 0)
    o = OracleCase(c)
    P = c.P.copy()
    t = 0.0
    steps = 0
    while t < t_end - 1e-15:
        dt = o.step(P, 1)[0]
        t += dt
        steps += 1
    V = c.layout.interior(P, 0)[:, 0, 0, :]
    return c, V, t, steps


@pytest.mark.parametrize("n", [200, 400, 800])
def test_sod_matches_exact(n):
    c, V, t, _ = run_sod(n, 0.1)
    g = c.meta["gamma"]
    ex = sample(c.meta["x"], t, 0.5, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, g)
    l1 = np.mean(np.abs(V[0] - ex[0]))
    assert l1 < 2.5 / n, l1          # first-order convergence of the density L1 error at discontinuities
    ps, us = star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, g)
    x = c.meta["x"]
    plateau = (x > 0.5 + us * t + 0.02) & (x < 0.5 + 1.6 * t)     # between the contact and the shock
    assert np.max(np.abs(V[4][plateau] - ps)) < 2e-3
    assert np.max(np.abs(V[1][plateau] - us)) < 5e-3


def test_sod_error_decreases():
    errs = []
    for n in (100, 200, 400):
        c, V, t, _ = run_sod(n, 0.1)
        ex = sample(c.meta["x"], t, 0.5, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, c.meta["gamma"])
        errs.append(np.mean(np.abs(V[0] - ex[0])))
    assert errs[0] > errs[1] > errs[2]
    assert errs[0] / errs[2] > 2.5


def test_sod_conservation():
    c = cases.sod(400, 0)
    o = OracleCase(c)
    P = c.P.copy()
    vol = c.layout.interior(c.geom.vol, 0)[0, 0, :]

    def totals(P):
        V = c.layout.interior(P, 0)[:, 0, 0, :]
        r, u, p = V[0], V[1], V[4]
        E = p / (c.meta["gamma"] - 1.0) + 0.5 * r * u * u
        return np.array([np.sum(r * vol), np.sum(r * u * vol), np.sum(E * vol)])

    t0 = totals(P)
    o.step(P, 60)            # t ~ 0.05: no wave has reached the ends
    t1 = totals(P)
    # mass and energy are conserved; momentum changes by the pressure difference of the two ends times t
    assert abs(t1[0] - t0[0]) < 1e-13 * t0[0]
    assert abs(t1[2] - t0[2]) < 1e-12 * t0[2]


def test_uniform_state_zero_residual():
    c = cases.taylor_green(12, 0)
    o = OracleCase(c)
    P = c.P.copy()
    P[0] = 1.2; P[1] = 0.0; P[2] = 0.0; P[3] = 0.0; P[4] = 101325.0
    R = o.residual(P)
    assert np.abs(c.layout.interior(R, 0)).max() == 0.0          # a gas at rest: every face flux cancels exactly
    P[1] = 3.0; P[2] = -2.0; P[3] = 1.0
    R = o.residual(P)
    area = (2 * np.pi / 12) ** 2
    H = 1004.5 * 101325.0 / (1.2 * 8314.46261815324 / 28.970418) + 0.5 * 14.0
    scale = (101325.0 + 1.2 * np.sqrt(14.0) * H) * area      # the largest face flux (energy)
    # moving: the periodic faces go through the record formulas (divides instead of reciprocals), ulps apart
    assert np.abs(c.layout.interior(R, 0)).max() <= 1e-15 * scale


def test_periodic_discrete_conservation():
    c = cases.taylor_green(12, 0)
    o = OracleCase(c)
    P = c.P.copy()
    R = o.residual(P)
    Ri = c.layout.interior(R, 0).reshape(5, -1)
    scale = np.abs(Ri).max(axis=1)
    # sum of the residual over the periodic box = 0 up to the rounding of the face fluxes
    assert np.all(np.abs(Ri.sum(axis=1)) < 1e-12 * scale * Ri.shape[1])


def test_rk3_linear_amplification():
    # constant-state advection: every residual is 0, the RK stages must leave the state bitwise unchanged
    c = cases.taylor_green(8, 0)
    o = OracleCase(c)
    P = c.P.copy()
    P[0] = 1.1; P[1] = 20.0; P[2] = 0.0; P[3] = 0.0; P[4] = 100000.0
    P0 = P.copy()
    o.step(P, 3)
    I = c.layout.interior(P, 0)
    I0 = c.layout.interior(P0, 0)
    assert np.abs(I - I0).max() <= 1e-12 * np.abs(I0).max()


def test_taylor_green_energy_decays():
    c = cases.taylor_green(16, 0)
    o = OracleCase(c)
    P = c.P.copy()
    vol = c.layout.interior(c.geom.vol, 0)

    def ke(P):
        V = c.layout.interior(P, 0)
        return np.sum(0.5 * V[0] * (V[1] ** 2 + V[2] ** 2 + V[3] ** 2) * vol)

    k0 = ke(P)
    o.step(P, 20)
    k1 = ke(P)
    assert k1 < k0 and k1 > 0.5 * k0   # classical Roe (Mco = 1) at M = 0.1 on 16^3 is dissipative
