"""The oracle's SLAU flux (riemann-solver = SLAU; the interior kernels' conv_physics.inc:126-167 and the record
kernels' riemann_SLAU, src/lib/numerics/fluxes/Lib_Riemann.f90:140-198), pinned without the reference.

- Consistency: for equal left and right states SLAU is the physical flux (the pressure weights b_l + b_r = 1, the
  mass flux rho u_n), which a dropped term or a wrong sign in the pressure or the mass flux breaks.
- The record variant computes its own total enthalpy: a uniform flow on a periodic box (every boundary face a
  connection record) has a zero residual, to rounding, only if the interior faces and the records agree.
- Sod's shock tube against Toro's exact solution: first-order convergence of the density's L1 error, the star-state
  plateau."""
import numpy as np
import pytest

from exact_riemann import sample, star_state
from mosegpu import cases
from oracle_bind import OracleCase


def run_sod(n, t_end, riemann="SLAU"):
    c = cases.sod(n, 0, riemann=riemann)
    o = OracleCase(c)
    P = c.P.copy()
    t = 0.0
    while t < t_end - 1e-15:
        t += o.step(P, 1)[0]
    return c, c.layout.interior(P, 0)[:, 0, 0, :], t


def test_slau_is_the_physical_flux_of_equal_states():
    c = cases.sod(8, 0, riemann="SLAU")
    o = OracleCase(c, nthreads=1)
    R = cases.R_AIR
    rng = np.random.Generator(np.random.PCG64(4))
    for _ in range(20):
        rho = rng.uniform(0.1, 3.0)
        u = rng.uniform(-600.0, 600.0, 3)
        p = rng.uniform(1e4, 3e5)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        q = np.array([rho, u[0], u[1], u[2], p])
        F = o.riemann(3, q, q, n)
        T = p / (rho * R)
        h0 = cases.CP_AIR * T + 0.5 * (u @ u)      # the constant-cp air table: h = cp T
        un = u @ n
        want = np.array([rho * un, rho * un * u[0] + p * n[0], rho * un * u[1] + p * n[1], rho * un * u[2] + p * n[2],
                         rho * un * h0])
        scale = np.array([abs(rho) * np.linalg.norm(u) + 1e-30] + [rho * (u @ u) + p] * 3 + [rho * np.linalg.norm(u) * h0])
        assert np.all(np.abs(F - want) <= 1e-13 * scale), (F, want)


def test_slau_uniform_flow_on_a_periodic_box_has_no_residual():
    c = cases.box_case((6, 5, 4), {f: "periodic" for f in range(1, 7)}, viscous=False, riemann="SLAU",
                       limiter="vanalbada")
    V = c.layout.interior(c.P, 0)
    V[0], V[1], V[2], V[3], V[4] = 1.2, 80.0, -30.0, 45.0, 1.1e5
    cases.fill_ghosts_nearest(c.layout, c.P)
    R = OracleCase(c, nthreads=1).residual(c.P.copy())
    A = c.layout.interior(R, 0)
    # fluxes of order rho |u| |u| area ~ 1e4 * 0.04; the residual of a uniform flow is rounding of those
    assert np.max(np.abs(A)) <= 1e-9, np.max(np.abs(A))


@pytest.mark.parametrize("n", [200, 400])
def test_slau_sod_matches_exact(n):
    c, V, t = run_sod(n, 0.1)
    g = c.meta["gamma"]
    ex = sample(c.meta["x"], t, 0.5, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, g)
    l1 = np.mean(np.abs(V[0] - ex[0]))
    assert l1 < 2.5 / n, l1
    ps, us = star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, g)
    x = c.meta["x"]
    plateau = (x > 0.5 + us * t + 0.02) & (x < 0.5 + 1.6 * t)
    assert np.max(np.abs(V[4][plateau] - ps)) < 3e-3
    assert np.max(np.abs(V[1][plateau] - us)) < 6e-3


def test_slau_sod_error_decreases():
    errs = []
    for n in (100, 200, 400):
        c, V, t = run_sod(n, 0.1)
        ex = sample(c.meta["x"], t, 0.5, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, c.meta["gamma"])
        errs.append(np.mean(np.abs(V[0] - ex[0])))
    assert errs[0] > errs[1] > errs[2] and errs[0] / errs[2] > 2.5, errs
