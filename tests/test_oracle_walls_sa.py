"""Pins of the oracle's wall, far-field and Spalart-Allmaras parts (CPU).

- A closed box of no-slip walls holding gas at rest has a zero residual.
  This covers BC_Symmetry_Eul, BC_Wall_Heat_dev and BC_Wall_Temperature_dev with T_wall = T.
- Isothermal walls at rest with T != T_wall conduct exactly 2 k(T_wall) (T - T_wall) / h per unit area.
  That is the wall-normal gradient over half a cell.
- A uniform free stream is a steady state of the far-field closures when the far-field data hold the stream's
  own stagnation state and static pressure.
  This covers BC_Outflow and BC_Subsonic_Inflow with the face-normal and the explicit flow direction.
- On a periodic shear layer the nu~ residual is minus the textbook SA source.
  The source is Spalart & Allmaras 1992 without ft2, with the SA-neg S~ branch.
  It is taken with and without the Patankar factor.
- Between two walls at rest the nu~ residual is the 1D diffusion stencil of (mu + rho nu~)/sigma grad nu~ plus the
  wall fluxes (nu~_wall = 0, the gradient over half a cell) minus the source.
  The source's nu~ gradient at the wall cells uses the negated ghost of Extrapolate_Wall.
"""
import numpy as np
import pytest

from mosegpu import cases
from mosegpu.core import BC_FARFIELD, BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL, FARFIELD_DTYPE, GC
from oracle_bind import OracleCase

R_AIR, CP_AIR = cases.R_AIR, cases.CP_AIR
GAM = CP_AIR / (CP_AIR - R_AIR)
# Spalart-Allmaras constants (the 1992 paper; ft2 omitted)
CB1, CB2, SIG, KAP, CW2, CW3, CV1, C2, C3 = 0.1355, 0.622, 2.0 / 3.0, 0.41, 0.3, 2.0, 7.1, 0.7, 0.9
CW1 = CB1 / KAP ** 2 + (1.0 + CB2) / SIG


def fill(case, rho, u, v, w, p, nt=None):
    for b in range(case.layout.nb):
        V = case.layout.interior(case.P, b)
        V[0], V[1], V[2], V[3], V[4] = rho, u, v, w, p
        if nt is not None:
            V[5] = nt
    cases.fill_ghosts_nearest(case.layout, case.P)


def table(t, T):
    Ti = np.floor(T).astype(int)
    return t[0, Ti] + (t[0, Ti + 1] - t[0, Ti]) * (T - Ti)


@pytest.mark.parametrize("typ", [BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL])
@pytest.mark.parametrize("rs", ["HLLC", "LMRoe", "AUSM+"])
def test_wall_box_at_rest(typ, rs):
    rho, p = 1.2, 1.0e5
    c = cases.box_case((5, 6, 4), {f: typ for f in range(1, 7)}, extent=(1.0, 1.2, 0.8), riemann=rs,
                       T_wall=p / (rho * R_AIR))
    fill(c, rho, 0.0, 0.0, 0.0, p)
    R = OracleCase(c).residual(c.P.copy())
    Ri = c.layout.interior(R, 0)
    area = 1.2 / 6 * 0.8 / 4
    assert np.max(np.abs(Ri)) <= 1e-12 * p * area


def test_isothermal_wall_conduction():
    rho, p, Tw = 1.1, 1.0e5, 250.0
    n = (3, 8, 3)
    ext = (0.3, 0.8, 0.3)
    faces = {1: "periodic", 2: "periodic", 3: BC_WALL_ISOTHERMAL, 4: BC_WALL_ISOTHERMAL, 5: "periodic", 6: "periodic"}
    c = cases.box_case(n, faces, extent=ext, T_wall=Tw)
    fill(c, rho, 0.0, 0.0, 0.0, p)
    R = c.layout.interior(OracleCase(c).residual(c.P.copy()), 0)
    T = p / (rho * R_AIR)
    h = ext[1] / n[1]
    area = (ext[0] / n[0]) * (ext[2] / n[2])
    k_w = table(c.tables.k, np.array(Tw))
    q = area * k_w * 2.0 * (T - Tw) / h            # leaves the cell through the wall face (R = outflow)
    assert np.allclose(R[4][:, 0, :], q, rtol=1e-12, atol=0)
    assert np.allclose(R[4][:, -1, :], q, rtol=1e-12, atol=0)
    assert np.max(np.abs(R[4][:, 1:-1, :])) <= 1e-12 * abs(q)
    assert np.max(np.abs(R[1:4])) <= 1e-12 * p * area      # no velocity, no stress


@pytest.mark.parametrize("alpha", [0.0, 0.3])
def test_farfield_uniform_stream_is_steady(alpha):
    rho, p, M = 1.15, 9.0e4, 0.3
    T = p / (rho * R_AIR)
    a = np.sqrt(GAM * R_AIR * T)
    U = M * a
    u, v = U * np.cos(alpha), U * np.sin(alpha)
    T0 = T * (1.0 + 0.5 * (GAM - 1.0) * M * M)
    ff = np.zeros(1, dtype=FARFIELD_DTYPE)
    ff["mach"] = 0.0; ff["t0"] = T0; ff["p0"] = p * (T0 / T) ** (GAM / (GAM - 1.0)); ff["pamb"] = p
    ff["rel_fac"] = 1.0; ff["ci"][0, 0] = 1.0; ff["alpha"] = alpha
    c = cases.box_case((6, 5, 4), {f: BC_FARFIELD for f in range(1, 7)}, extent=(1.0, 0.9, 0.7), farfield=ff,
                       viscous=False)
    fill(c, rho, u, v, 0.0, p)
    R = c.layout.interior(OracleCase(c).residual(c.P.copy()), 0)
    area = 0.9 / 5 * 0.7 / 4
    scale = [rho * U * area, (rho * U * U + p) * area, (rho * U * U + p) * area, p * area,
             rho * U * (CP_AIR * T0) * area]
    for s in range(5):
        assert np.max(np.abs(R[s])) <= 1e-10 * scale[s], (s, np.max(np.abs(R[s])))


def sa_source(rho, mu, rnt, Om, gn2, d, vol, dt=None):
    """Textbook SA source per unit volume times rho (the sign of a production), times the volume."""
    nit = rnt / rho
    chi = rnt / mu
    fv1 = chi ** 3 / (chi ** 3 + CV1 ** 3)
    fv2 = 1.0 - chi / (1.0 + chi * fv1)
    Sb = nit / (KAP ** 2 * d ** 2) * fv2
    St = np.where(Sb >= -C2 * Om, Om + Sb, Om + Om * (C2 ** 2 * Om + C3 * Sb) / ((C3 - 2 * C2) * Om - Sb))
    St = np.where(St == 0.0, 1e-40, St)
    r = np.minimum(nit / (St * KAP ** 2 * d ** 2), 10.0)
    g = r + CW2 * (r ** 6 - r)
    fw = g * ((1.0 + CW3 ** 6) / (g ** 6 + CW3 ** 6)) ** (1.0 / 6.0)
    src = rho * (CB1 * St * nit + CB2 / SIG * gn2 - CW1 * fw * (nit / d) ** 2) * vol
    if dt is not None:
        src = src / (1.0 + dt * 2.0 * CW1 * fw * nit / d ** 2)
    return src


@pytest.mark.parametrize("pimp", [0, 1])
def test_sa_source_on_periodic_shear(pimp):
    n = (5, 12, 3)
    ext = (0.5, 1.2, 0.3)
    c = cases.box_case(n, {f: "periodic" for f in range(1, 7)}, extent=ext, nrans=1, point_implicit=pimp)
    rho, p, U = 1.2, 1.0e5, 30.0
    hy = ext[1] / n[1]
    yc = (np.arange(n[1]) + 0.5) * hy
    u = U * np.sin(2 * np.pi * yc / ext[1])
    T = p / (rho * R_AIR)
    mu = table(c.tables.mu, np.array(T))
    rnt = 40.0 * mu                                   # chi = 40
    fill(c, rho, u[None, :, None], 0.0, 0.0, p, rnt)
    d = 0.05 + yc                                     # any positive wall distance
    c.layout.interior(c.geom.yn, 0)[...] = d[None, :, None]
    dt = 3e-4
    R = c.layout.interior(OracleCase(c).residual(c.P.copy(), dt), 0)
    Om = np.abs((np.roll(u, -1) - np.roll(u, 1)) / (2.0 * hy))
    vol = np.prod(ext) / np.prod(n)
    src = sa_source(rho, mu, rnt, Om, 0.0, d, vol, dt if pimp else None)
    got = R[5][0, :, 0]
    assert np.allclose(got, -src, rtol=1e-12, atol=1e-12 * np.max(np.abs(src)))
    assert np.allclose(R[5], R[5][:1, :, :1], rtol=1e-12, atol=1e-12 * np.max(np.abs(src)))   # x, z invariance


def test_sa_between_walls_at_rest():
    n = (3, 7, 3)
    ext = (0.3, 0.7, 0.3)
    faces = {1: "periodic", 2: "periodic", 3: BC_WALL_ADIABATIC, 4: BC_WALL_ADIABATIC, 5: "periodic", 6: "periodic"}
    c = cases.box_case(n, faces, extent=ext, nrans=1, point_implicit=0)
    rho, p = 1.2, 1.0e5
    T = p / (rho * R_AIR)
    mu = table(c.tables.mu, np.array(T))
    rnt = mu * np.array([20.0, 30.0, 45.0, 55.0, 50.0, 35.0, 25.0])
    fill(c, rho, 0.0, 0.0, 0.0, p, rnt[None, :, None])
    hy = ext[1] / n[1]
    yc = (np.arange(n[1]) + 0.5) * hy
    d = np.minimum(yc, ext[1] - yc)
    c.layout.interior(c.geom.yn, 0)[...] = d[None, :, None]
    R = c.layout.interior(OracleCase(c).residual(c.P.copy()), 0)[5][0, :, 0]
    area = (ext[0] / n[0]) * (ext[2] / n[2])
    vol = area * hy
    nt = rnt / rho
    # interior faces j+1/2: D = area (mu + rho nu~_face) / sigma * d(nu~)/dy, R = -D(j+1/2) + D(j-1/2)
    D = area * (mu + 0.5 * (rnt[1:] + rnt[:-1])) * ((nt[1:] - nt[:-1]) / hy) / SIG
    exp = np.zeros(n[1])
    exp[:-1] -= D
    exp[1:] += D
    # the wall faces: nu~ = 0 on the wall, half a cell away
    exp[0] += area * mu * (2.0 * nt[0] / hy) / SIG
    exp[-1] += area * mu * (2.0 * nt[-1] / hy) / SIG
    # the source, with the negated wall ghost in the nu~ gradient
    ntg = np.concatenate([[-nt[0]], nt, [-nt[-1]]])
    gn = (ntg[2:] - ntg[:-2]) / (2.0 * hy)
    exp -= sa_source(rho, mu, rnt, 0.0, gn ** 2, d, vol)
    assert np.allclose(R, exp, rtol=1e-11, atol=1e-12 * np.max(np.abs(exp)))


def lattice_field(c):
    """A smooth state over the lattice's cell centres (the same function for any block split)."""
    U = c.meta["U"]
    for b, nd in enumerate(c.nodes):
        ctr = 0.125 * (nd[:-1, :-1, :-1] + nd[1:, :-1, :-1] + nd[:-1, 1:, :-1] + nd[:-1, :-1, 1:] + nd[1:, 1:, :-1]
                       + nd[1:, :-1, 1:] + nd[:-1, 1:, 1:] + nd[1:, 1:, 1:])
        x, y, z = ctr[..., 0], ctr[..., 1], ctr[..., 2]
        V = c.layout.interior(c.P, b)
        V[0] = c.meta["rho_inf"] * (1.0 + 0.02 * np.sin(3 * x) * np.cos(2 * z))
        V[1] = U * np.tanh(y / 0.01) * (1.0 + 0.05 * np.cos(4 * x))
        V[2] = 0.02 * U * np.sin(5 * x) * np.sin(7 * y)
        V[3] = 0.03 * U * np.cos(3 * z + y)
        V[4] = 101325.0 * (1.0 + 0.01 * np.cos(2 * x + 3 * z))
        V[5] = c.meta["mu_inf"] * (2.0 + np.sin(6 * x + 4 * y + z))
    cases.fill_ghosts_nearest(c.layout, c.P)


def global_cells(c, arr):
    """The interior of a lattice arena as one (nvar, Nz, Ny, Nx) array."""
    wx, wy, wz = c.meta["widths"]
    out = np.zeros((arr.shape[0], sum(wz), sum(wy), sum(wx)))
    b = 0
    oz = 0
    for cz in wz:
        oy = 0
        for cy in wy:
            ox = 0
            for cx in wx:
                out[:, oz:oz + cz, oy:oy + cy, ox:ox + cx] = c.layout.interior(arr, b)
                b += 1
                ox += cx
            oy += cy
        oz += cz
    return out


@pytest.mark.parametrize("wall", [BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL])
def test_lattice_equals_one_block(wall):
    """Block-to-block connections (halo copy + connection records) reproduce the single-block interior faces to
    round-off: the residual and three SSP-RK3 steps of a 2x2x2 lattice equal those of the same grid as one block."""
    kw = dict(wall=wall, T_wall=310.0, extent=(0.4, 0.2, 0.3), nsteps=3)
    one = cases.channel_lattice(nb=(1, 1, 1), widths=[[9], [8], [7]], **kw)
    lat = cases.channel_lattice(nb=(2, 2, 2), widths=[[4, 5], [5, 3], [3, 4]], **kw)
    for c in (one, lat):
        lattice_field(c)
    R1 = global_cells(one, OracleCase(one).residual(one.P.copy(), 1e-6))
    R8 = global_cells(lat, OracleCase(lat).residual(lat.P.copy(), 1e-6))
    for s in range(6):
        scale = np.max(np.abs(R1[s]))
        assert np.max(np.abs(R8[s] - R1[s])) <= 1e-11 * scale, s
    P1, P8 = one.P.copy(), lat.P.copy()
    OracleCase(one).step(P1, 3)
    OracleCase(lat).step(P8, 3)
    G1, G8 = global_cells(one, P1), global_cells(lat, P8)
    for s in range(6):
        scale = np.max(np.abs(G1[s]))
        assert np.max(np.abs(G8[s] - G1[s])) <= 1e-12 * scale, s


@pytest.mark.parametrize("rs", ["HLLC", "LMRoe", "AUSM+"])
def test_farfield_supersonic_stream_is_steady(rs):
    """BC_Supersonic_Inflow takes (Mach, p, T) of the far-field row as the static inflow state (MOSE_open 405 /
    Mach > 1) and BC_Outflow extrapolates a supersonic exit: a uniform M = 2 stream at an angle is steady."""
    rho, p, M, al, be = 0.9, 5.0e4, 2.0, 0.2, -0.1
    T = p / (rho * R_AIR)
    U = M * np.sqrt(GAM * R_AIR * T)
    u, v, w = U * np.cos(al) * np.cos(be), U * np.sin(al) * np.cos(be), U * np.sin(be)
    ff = np.zeros(1, dtype=FARFIELD_DTYPE)
    ff["mach"] = M; ff["t0"] = T; ff["p0"] = p; ff["pamb"] = p; ff["rel_fac"] = 1.0; ff["ci"][0, 0] = 1.0
    ff["alpha"] = al; ff["beta"] = be
    c = cases.box_case((6, 5, 4), {f: BC_FARFIELD for f in range(1, 7)}, extent=(1.0, 0.9, 0.7), farfield=ff,
                       viscous=False, riemann=rs)
    fill(c, rho, u, v, w, p)
    R = c.layout.interior(OracleCase(c).residual(c.P.copy()), 0)
    area = 0.9 / 5 * 0.7 / 4
    scale = rho * U * U * area + p * area
    assert np.max(np.abs(R)) <= 1e-11 * scale * (1.0 + CP_AIR * T / U ** 2)
