"""Pins of the oracle's time step, dt and viscous stress against closed forms (CPU).

SSP-RK3 (Lib_RK.f90:6-9, Mod_Explicit.f90:159-279): a density sine mode advected at uniform velocity and pressure
through first-order upwind HLLC.  HLLC resolves a contact exactly, so the mass flux of a face is rho_left u and the
semi-discrete system is linear, d rho_i / dt = -(u / dx) (rho_i - rho_(i-1)); u and p stay uniform.  A Fourier mode
exp(i k x) then gains R(z) = 1 + z + z^2/2 + z^3/6 per step, z = -nu (1 - exp(-i k dx)), nu = u dt / dx, with each
step's own dt.  The product of the R(z_n) is checked to 1e-12 (amplitude and phase); the stage coefficients 1/4
and 2/3 written as 1/2 and 1/2 miss it by 1e-4.

dt (dt_blk.inc:157-191): on a uniform state the CFL limit is min over directions of cfl dl_d / (|u . m_d / |m_d|| + a)
and the viscous limit rho dl_d^2 vnn / mu; both closed forms on an anisotropic box, in the regime where each wins.

Viscous stress (diff_physics.inc:59-161): for a quadratic velocity field on any affine mesh the face gradients
(normal two-cell difference, transverse four-cell averages, the face-averaged metric) are exact at the face centre,
and the face fluxes area (tau . n) sum to vol div(tau) exactly.  With uniform rho and p (uniform T, constant mu) the
viscous part of the momentum residual is then -vol (mu lap(u) + (mu + lambda) grad(div u)), lambda = -2/3 mu.  The
viscous part is the residual of the viscous scheme minus that of the same scheme without viscosity.  lambda = +2/3 mu,
a dropped transpose term or a transposed metric misses it by O(1).
"""
import numpy as np
import pytest

from mosegpu import cases
from mosegpu.core import BC_EXTRAPOLATION, BC_NONE, GC
from oracle_bind import OracleCase

GAMMA = cases.CP_AIR / (cases.CP_AIR - cases.R_AIR)


def rk3_factor(z, c2=0.25, c3=2.0 / 3.0):
    """One step of U_k = U0 + (U_(k-1) - U0 + dt L U_(k-1)) c_k, c = (1, c2, c3), on dU/dt = (z/dt) U."""
    u1 = 1.0 + z
    u2 = 1.0 + (u1 - 1.0 + z * u1) * c2
    return 1.0 + (u2 - 1.0 + z * u2) * c3


def advection_case(n, u, p, rho0, eps, m):
    c = cases.box_case((n, 1, 1), {1: "periodic", 2: "periodic", 3: BC_NONE, 4: BC_NONE, 5: BC_NONE, 6: BC_NONE},
                       extent=(1.0, 0.1, 0.1), viscous=False, ndir=1, limiter="first-order", riemann="HLLC", cfl=0.5)
    x = (np.arange(n) + 0.5) / n
    V = c.layout.interior(c.P, 0)
    V[0] = rho0 + eps * np.sin(2 * np.pi * m * x)
    V[1] = u
    V[4] = p
    cases.fill_ghosts_nearest(c.layout, c.P)
    return c, x


def fourier(rho, x, m):
    return 2.0 / rho.size * np.sum(rho * np.exp(-2j * np.pi * m * x))


def test_ssp_rk3_amplification_of_an_advected_mode():
    n, m, u, p, rho0, eps, nsteps = 32, 1, 120.0, 1.0e5, 1.0, 0.1, 40
    c, x = advection_case(n, u, p, rho0, eps, m)
    o = OracleCase(c)
    P = c.P.copy()
    c0 = fourier(c.layout.interior(P, 0)[0, 0, 0], x, m)
    dx, k = 1.0 / n, 2 * np.pi * m
    pred = c0
    wrong = c0
    for _ in range(nsteps):
        dt = o.step(P, 1)[0]
        z = -(u * dt / dx) * (1.0 - np.exp(-1j * k * dx))
        pred *= rk3_factor(z)
        wrong *= rk3_factor(z, 0.5, 0.5)
    V = c.layout.interior(P, 0)[:, 0, 0]
    got = fourier(V[0], x, m)
    assert abs(got - pred) <= 1e-12 * abs(c0), abs(got - pred) / abs(c0)
    assert abs(wrong - pred) > 1e-5 * abs(c0)          # the check can tell the coefficients apart
    # the mode is the only content: the other modes stay at rounding level, u and p stay uniform
    rest = V[0] - rho0 - (got.real * np.cos(k * x) - got.imag * np.sin(k * x))
    assert np.max(np.abs(rest)) <= 1e-12
    assert np.max(np.abs(V[1] - u)) <= 1e-12 * u and np.max(np.abs(V[4] - p)) <= 1e-12 * p


def test_dt_cfl_closed_form():
    dims, ext = (6, 5, 4), (0.6, 1.0, 2.0)            # dl = 0.1, 0.2, 0.5
    c = cases.box_case(dims, {f: "periodic" for f in range(1, 7)}, extent=ext, viscous=False, cfl=0.7)
    rho, vel, p = 1.3, np.array([310.0, -150.0, 40.0]), 9.0e4
    V = c.layout.interior(c.P, 0)
    V[0], V[1], V[2], V[3], V[4] = rho, vel[0], vel[1], vel[2], p
    cases.fill_ghosts_nearest(c.layout, c.P)
    a = np.sqrt(GAMMA * p / rho)
    dl = np.array(ext) / np.array(dims)
    want = min(0.7 * dl[d] / (abs(vel[d]) + a) for d in range(3))
    got = OracleCase(c).dt(c.P)
    assert abs(got - want) <= 1e-14 * want, (got, want)


def test_dt_viscous_closed_form():
    dims, ext = (6, 5, 4), (0.06, 0.1, 0.2)           # dl = 0.01, 0.02, 0.05
    mu = 20.0
    c = cases.box_case(dims, {f: "periodic" for f in range(1, 7)}, extent=ext, viscous=True,
                       mu_of_T=lambda T: mu, cfl=0.7, vnn=0.4)
    rho, vel, p = 1.3, np.array([30.0, -15.0, 4.0]), 9.0e4
    V = c.layout.interior(c.P, 0)
    V[0], V[1], V[2], V[3], V[4] = rho, vel[0], vel[1], vel[2], p
    cases.fill_ghosts_nearest(c.layout, c.P)
    dl = np.array(ext) / np.array(dims)
    a = np.sqrt(GAMMA * p / rho)
    cfl_lim = min(0.7 * dl[d] / (abs(vel[d]) + a) for d in range(3))
    vnn_lim = min(rho * dl[d] ** 2 * 0.4 / mu for d in range(3))
    assert vnn_lim < 0.5 * cfl_lim                    # the viscous limit wins here
    got = OracleCase(c).dt(c.P)
    assert abs(got - vnn_lim) <= 1e-14 * vnn_lim, (got, vnn_lim)


J_SKEW = np.array([[0.031, 0.004, -0.006],
                   [0.003, 0.027, 0.002],
                   [-0.002, 0.005, 0.036]])


def quadratic_field(X, A, B):
    """u_a = A[a] . X + X . B[a] X (a = 0..2) at points X (..., 3)."""
    lin = np.einsum("ab,...b->...a", A, X)
    quad = np.einsum("...i,aij,...j->...a", X, B, X)
    return lin + quad


def viscous_force(A, B, mu):
    """mu lap(u) + (mu + lambda) grad(div u) of quadratic_field (constant), lambda = -2/3 mu."""
    Bs = B + np.transpose(B, (0, 2, 1))            # Hessian of u_a
    lap = np.array([np.trace(Bs[a]) for a in range(3)])
    graddiv = np.array([sum(Bs[b][b, a] for b in range(3)) for a in range(3)])
    return mu * lap + (mu - 2.0 / 3.0 * mu) * graddiv


@pytest.mark.parametrize("mesh", ["cartesian", "skewed"])
def test_viscous_stress_of_a_quadratic_field(mesh):
    dims = (9, 8, 7)
    n1, n2, n3 = dims
    J = np.diag([0.03, 0.025, 0.035]) if mesh == "cartesian" else J_SKEW
    k, j, i = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    nodes = np.stack([i, j, k], axis=-1).astype(np.float64) @ J.T
    mu = 2.0
    rng = np.random.Generator(np.random.PCG64(3))
    A = rng.uniform(-400, 400, (3, 3))
    B = rng.uniform(-2000, 2000, (3, 3, 3))
    out = {}
    for visc in (False, True):
        c = cases.box_case(dims, {f: BC_EXTRAPOLATION for f in range(1, 7)}, viscous=visc, nodes=nodes,
                           mu_of_T=(lambda T: mu) if visc else None, riemann="HLLC", limiter="minmod", ndir=3)
        # cell centres = the mean of the eight nodes (the centroid of a parallelepiped)
        ctr = 0.125 * (nodes[:-1, :-1, :-1] + nodes[:-1, :-1, 1:] + nodes[:-1, 1:, :-1] + nodes[:-1, 1:, 1:] +
                       nodes[1:, :-1, :-1] + nodes[1:, :-1, 1:] + nodes[1:, 1:, :-1] + nodes[1:, 1:, 1:])
        Vb = c.layout.block_view(c.P, 0)
        # the quadratic field on every padded cell, ghosts included (centres extended affinely)
        kk, jj, ii = np.meshgrid(np.arange(-GC, n3 + GC), np.arange(-GC, n2 + GC), np.arange(-GC, n1 + GC),
                                 indexing="ij")
        Xp = (np.stack([ii, jj, kk], axis=-1) + 0.5) @ J.T
        U = quadratic_field(Xp, A, B)
        Vb[0] = 1.2
        Vb[1], Vb[2], Vb[3] = U[..., 0], U[..., 1], U[..., 2]
        Vb[4] = 1.0e5
        assert np.allclose(Xp[GC:-GC, GC:-GC, GC:-GC], ctr)
        o = OracleCase(c)
        R = o.residual(c.P.copy())
        out[visc] = c.layout.interior(R, 0)
        vol = c.layout.interior(c.geom.vol.reshape(1, -1), 0)[0]
    Rv = out[True] - out[False]                        # the viscous part: -sum of the viscous face fluxes
    want = -viscous_force(A, B, mu)
    inner = (slice(2, -2),) * 3                         # cells whose faces see no boundary record or ghost
    for a in range(3):
        got = Rv[1 + a][inner] / vol[inner]
        scale = np.abs(want).max()
        assert np.max(np.abs(got - want[a])) <= 1e-8 * scale, (a, np.max(np.abs(got - want[a])) / scale)
    # lambda = +2/3 mu would give a different force by (4/3) mu grad(div u)
    Bs = B + np.transpose(B, (0, 2, 1))
    graddiv = np.array([sum(Bs[b][b, a] for b in range(3)) for a in range(3)])
    assert np.max(np.abs(4.0 / 3.0 * mu * graddiv)) > 1e-3 * np.abs(want).max()
