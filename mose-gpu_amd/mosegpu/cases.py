"""Synthetic cases of the residual path: thermo tables, meshes, boundary records and initial fields.

The cases follow the reference's case format semantics (test/1D/Sod79/INPUT: a constant-cp air table with
h = cp T, FLINT's T = 0..10000 K rows; MOSE boundary types 400 -> 11, 201 -> 1 periodic), built in memory:
  sod(n)            1D shock tube, n x 1 x 1 on x in [0, 1], minmod + HLLC, SSP-RK3, CFL 0.5, transmissive ends
  taylor_green(n)   3D periodic n^3 on [0, 2 pi]^3, M = 0.1, Re = 1600, Roe (LMRoe with Mco = 1) + van Albada,
                    Sutherland viscosity tabulated per Kelvin, Pr = 0.72, SSP-RK3 at CFL 0.4, seeded velocity
                    perturbation (at CFL 0.5 this low-Mach Roe flux at M = 0.1 is unstable: the pressure minimum
                    collapses after about 40 steps at n = 32 and 120 steps at n = 54; CFL 0.4 holds through 500
                    steps at n = 32 and n = 64)
  channel_lattice() a lattice of blocks of seeded uneven widths with one-to-one interfaces (MOSE 201 -> 1 without a
                    period), a no-slip wall at y = 0 (301 -> 5 adiabatic or 302 -> 6 isothermal), far field (407 -> 4)
                    on the other faces, Spalart-Allmaras, M = 0.3, Re = 1e6
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .core import (BC_CONNECTION, BC_EXTRAPOLATION, BC_NONE, BC_FARFIELD, BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL,
                   FARFIELD_DTYPE, GC, INTEGRATION, LIM, NTAB, RECORD_DTYPE, RS, Geometry, Layout, Scheme, ThermoTables,
                   metrics_build, records_tangent, wall_distance)

RUNIV = 8314.46261815324
WM_AIR = 28.970418          # test/1D/Sod79/INPUT/phase.txt
CP_AIR = 1004.5             # test/1D/Sod79/INPUT/thermo.dat
R_AIR = RUNIV / WM_AIR


def air_tables(mu_of_T=None, prandtl: float | None = None) -> ThermoTables:
    """The constant-cp air table of test/1D/Sod79/INPUT/thermo.dat on T = 0..10000 K (FLINT fills row 0 with
    cp(1) and h = 0, Load_ThermoTransport.f90:92-97), plus optional transport rows mu(T) and k = mu cp / Pr."""
    T = np.arange(NTAB, dtype=np.float64)
    cp = np.full(NTAB, CP_AIR)
    h = CP_AIR * T
    h[0] = 0.0
    mu = k = None
    if mu_of_T is not None:
        mu = np.array([mu_of_T(t) for t in T[1:]], dtype=np.float64)
        mu = np.concatenate([[mu[0]], mu])            # row 0 = row 1 (Load_ThermoTransport.f90:184-187)
        k = mu * CP_AIR / (prandtl if prandtl else 0.72)
    return ThermoTables([WM_AIR], cp[None, :], h[None, :], None if mu is None else mu[None, :],
                        None if k is None else k[None, :], names=["air"])


def sutherland(mu0: float, T0: float, S: float = 110.4):
    def f(T):
        return mu0 * (T / T0) ** 1.5 * (T0 + S) / (T + S)
    return f


def box_nodes(n1, n2, n3, lo, hi):
    """Nodes of a Cartesian block, shape (n3+1, n2+1, n1+1, 3), i fastest."""
    xs = [lo[a] + (hi[a] - lo[a]) * np.arange(n + 1, dtype=np.float64) / n for a, n in enumerate((n1, n2, n3))]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return np.stack([X, Y, Z], axis=-1)


def face_cells(n1, n2, n3, f):
    """(i, j, k) of the boundary cells of face f (1..6), i fastest."""
    if f in (1, 2):
        k, j = np.meshgrid(np.arange(1, n3 + 1), np.arange(1, n2 + 1), indexing="ij")
        i = np.full_like(j, 1 if f == 1 else n1)
    elif f in (3, 4):
        k, i = np.meshgrid(np.arange(1, n3 + 1), np.arange(1, n1 + 1), indexing="ij")
        j = np.full_like(i, 1 if f == 3 else n2)
    else:
        j, i = np.meshgrid(np.arange(1, n2 + 1), np.arange(1, n1 + 1), indexing="ij")
        k = np.full_like(i, 1 if f == 5 else n3)
    return i.reshape(-1), j.reshape(-1), k.reshape(-1)


def block_records(b, dims, face_types, donors=None):
    """Records of block b: face_types[f] = type; donors[f] = (bs, fs, map) for connections, where map turns the
    receiver cell (i, j, k) into the donor cell."""
    n1, n2, n3 = dims
    out = []
    for f in range(1, 7):
        i, j, k = face_cells(n1, n2, n3, f)
        r = np.zeros(i.size, dtype=RECORD_DTYPE)
        r["b"] = b; r["i"] = i; r["j"] = j; r["k"] = k; r["f"] = f; r["typ"] = face_types[f]
        r["d11"] = 1; r["d22"] = 1
        if face_types[f] == BC_CONNECTION:
            bs, fs, fmap, periodic = donors[f]
            si, sj, sk = fmap(i, j, k)
            r["bs"] = bs; r["fs"] = fs; r["is_"] = si; r["js"] = sj; r["ks"] = sk; r["periodic"] = int(periodic)
        out.append(r)
    return np.concatenate(out)


def periodic_box_records(n1, n2, n3):
    """All six faces of one block connected to the opposite face (MOSE 201)."""
    donors = {
        1: (0, 2, lambda i, j, k: (np.full_like(i, n1), j, k), True),
        2: (0, 1, lambda i, j, k: (np.full_like(i, 1), j, k), True),
        3: (0, 4, lambda i, j, k: (i, np.full_like(j, n2), k), True),
        4: (0, 3, lambda i, j, k: (i, np.full_like(j, 1), k), True),
        5: (0, 6, lambda i, j, k: (i, j, np.full_like(k, n3)), True),
        6: (0, 5, lambda i, j, k: (i, j, np.full_like(k, 1)), True),
    }
    return block_records(0, (n1, n2, n3), {f: BC_CONNECTION for f in range(1, 7)}, donors)


def fill_ghosts_nearest(layout: Layout, P: np.ndarray):
    """Ghost rows (never read before the first ghost fill) set to the nearest interior cell."""
    for b in range(layout.nb):
        v = layout.block_view(P, b)
        for ax in (1, 2, 3):
            n = v.shape[ax] - 2 * GC
            for g in range(GC):
                lo = [slice(None)] * 4; src = [slice(None)] * 4
                lo[ax] = g; src[ax] = GC
                v[tuple(lo)] = v[tuple(src)]
                hi = [slice(None)] * 4; srh = [slice(None)] * 4
                hi[ax] = GC + n + g; srh[ax] = GC + n - 1
                v[tuple(hi)] = v[tuple(srh)]


@dataclass
class Case:
    name: str
    tables: ThermoTables
    scheme: Scheme
    layout: Layout
    nodes: list
    records: np.ndarray
    geom: Geometry
    P: np.ndarray                    # initial state [nprim, NC]
    nsteps: int
    meta: dict = field(default_factory=dict)
    farfield: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=FARFIELD_DTYPE))

    @property
    def nprim(self):
        return self.scheme.nsc + 4 + self.scheme.nrans

    def interior_cells(self):
        return sum(n1 * n2 * n3 for n1, n2, n3 in self.layout.dims)


def make_scheme(nsc=1, nrans=0, limiter="minmod", riemann="HLLC", viscous=False, n_rk=3, ndir=3, cfl=0.5,
                vnn=0.5, mco=1.0, prandtl=0.0, sc=0.5, sct=0.9, prt=0.9, point_implicit=1, guard=1,
                integration="cons", irs_beta=None, irs_njacobi=2) -> Scheme:
    """integration: the state update's variables, "cons" (Newstate_Conservative_blk) or "prim"
    (Newstate_Primitive_blk), as [MOSE-Numerics] integration-variables (Lib_Newstate.f90:43-60).
    irs_beta: implicit residual smoothing on with this coefficient (IRS_apply, Lib_IRS.f90:53-96), None = off;
    irs_njacobi: its Jacobi sweeps per direction."""
    s = Scheme()
    s.nsc = nsc; s.nrans = nrans; s.limiter = LIM[limiter]; s.riemann = RS[riemann]; s.viscous = int(viscous)
    s.n_rk = n_rk; s.ndir = ndir; s.point_implicit = point_implicit; s.thermo_guard = guard
    s.integration = INTEGRATION[integration]
    s.irs = 0 if irs_beta is None else 1
    s.irs_beta = 0.0 if irs_beta is None else float(irs_beta)
    s.irs_njacobi = irs_njacobi
    s.mco = mco; s.cfl = cfl; s.vnn = vnn; s.prandtl = prandtl; s.sc = sc; s.sct = sct; s.prt = prt
    return s


def sod(n: int = 1000, nsteps: int = 200, cfl: float = 0.5, limiter="minmod", riemann="HLLC") -> Case:
    """1D inviscid shock tube: n x 1 x 1 cells on x in [0, 1], y in [0, 0.01], z in [-5e-4, 5e-4] (the extents of
    test/1D/Sod79/MESH/mesh_x4.tec), (rho, u, p) = (1, 0, 1) left of x = 0.5 and (0.125, 0, 0.1) right of it,
    extrapolation (MOSE 400) on every face, SSP-RK3."""
    tables = air_tables()
    layout = Layout([(n, 1, 1)])
    nodes = [box_nodes(n, 1, 1, (0.0, 0.0, -5e-4), (1.0, 0.01, 5e-4))]
    rec = block_records(0, (n, 1, 1), {f: BC_EXTRAPOLATION for f in range(1, 7)})
    geom = metrics_build(layout, nodes, rec, ndir=1)
    sch = make_scheme(limiter=limiter, riemann=riemann, viscous=False, ndir=1, cfl=cfl)
    P = np.zeros((5, layout.NC))
    xc = (np.arange(n) + 0.5) / n
    left = xc < 0.5
    V = layout.interior(P, 0)          # (5, 1, 1, n)
    V[0, 0, 0, :] = np.where(left, 1.0, 0.125)
    V[4, 0, 0, :] = np.where(left, 1.0, 0.1)
    fill_ghosts_nearest(layout, P)
    return Case("sod", tables, sch, layout, nodes, rec, geom, P, nsteps,
                meta={"x": xc, "gamma": CP_AIR / (CP_AIR - R_AIR)})


def taylor_green(n: int = 64, nsteps: int = 100, mach: float = 0.1, re: float = 1600.0, T0: float = 300.0,
                 p0: float = 101325.0, amp: float = 1e-3, seed: int = 2024, cfl: float = 0.4, vnn: float = 0.5,
                 dims=None) -> Case:
    """3D compressible Taylor-Green vortex on the periodic box [0, 2 pi]^3 (L = 1), M = V0/a0, Re = rho0 V0 L / mu0,
    isothermal start rho = p / (R T0), plus a seeded uniform velocity perturbation of amplitude amp * V0."""
    dims = dims or (n, n, n)
    n1, n2, n3 = dims
    gam = CP_AIR / (CP_AIR - R_AIR)
    a0 = np.sqrt(gam * R_AIR * T0)
    V0 = mach * a0
    rho0 = p0 / (R_AIR * T0)
    mu0 = rho0 * V0 * 1.0 / re
    tables = air_tables(sutherland(mu0, T0), prandtl=0.72)
    layout = Layout([dims])
    L = 2.0 * np.pi
    nodes = [box_nodes(n1, n2, n3, (0.0, 0.0, 0.0), (L, L, L))]
    rec = periodic_box_records(n1, n2, n3)
    geom = metrics_build(layout, nodes, rec, ndir=3)
    sch = make_scheme(limiter="vanalbada", riemann="LMRoe", mco=1.0, viscous=True, ndir=3, cfl=cfl, vnn=vnn,
                      prandtl=0.72)
    P = np.zeros((5, layout.NC))
    x = (np.arange(n1) + 0.5) * (L / n1)
    y = (np.arange(n2) + 0.5) * (L / n2)
    z = (np.arange(n3) + 0.5) * (L / n3)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rng = np.random.Generator(np.random.PCG64(seed))
    pert = rng.uniform(-1.0, 1.0, size=(3,) + X.shape)
    u = V0 * np.sin(X) * np.cos(Y) * np.cos(Z) + amp * V0 * pert[0]
    v = -V0 * np.cos(X) * np.sin(Y) * np.cos(Z) + amp * V0 * pert[1]
    w = amp * V0 * pert[2]
    p = p0 + rho0 * V0 ** 2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    rho = p / (R_AIR * T0)
    V = layout.interior(P, 0)
    V[0], V[1], V[2], V[3], V[4] = rho, u, v, w, p
    fill_ghosts_nearest(layout, P)
    return Case("taylor_green", tables, sch, layout, nodes, rec, geom, P, nsteps,
                meta={"V0": V0, "rho0": rho0, "mu0": mu0, "a0": a0})


def channel_lattice(nb=(4, 4, 4), wmin: int = 24, wmax: int = 96, seed: int = 7, nsteps: int = 200, mach: float = 0.3,
                    re: float = 1e6, T_inf: float = 300.0, p_inf: float = 101325.0, wall: int = BC_WALL_ADIABATIC,
                    T_wall: float = 300.0, extent=(1.0, 0.3, 0.3), stretch: float = 3.0, riemann="HLLC",
                    limiter="vanalbada", cfl: float = 0.5, vnn: float = 0.5, point_implicit: int = 1,
                    widths=None) -> Case:
    """Channel-like flow over a no-slip wall on a lattice of nb[0] x nb[1] x nb[2] blocks.  The per-axis block widths
    are seeded integers in [wmin, wmax] (widths overrides them); block (a, b, c) has w_x[a] x w_y[b] x w_z[c] cells.
    Interfaces are one-to-one connections, the y = 0 face is a wall, every other outer face is far field with the
    free stream's stagnation state (inflow) and its static pressure (outflow).  The grid is uniform in x and z and
    stretched towards the wall in y.  State: free stream at M = mach with a tanh boundary layer of thickness
    0.02 extent[1], a 1e-3 seeded velocity perturbation, and rho nu~ = rho nu U(1, 5) (seeded).  Spalart-Allmaras,
    SSP-RK3, Pr = 0.72, Re = rho U L / mu with L = 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if widths is None:
        widths = [rng.integers(wmin, wmax + 1, size=n).tolist() for n in nb]
    wx, wy, wz = (list(map(int, w)) for w in widths)
    gam = CP_AIR / (CP_AIR - R_AIR)
    a_inf = np.sqrt(gam * R_AIR * T_inf)
    U = mach * a_inf
    rho_inf = p_inf / (R_AIR * T_inf)
    mu_inf = rho_inf * U * 1.0 / re
    tables = air_tables(sutherland(mu_inf, T_inf), prandtl=0.72)
    # global node lines
    Nx, Ny, Nz = sum(wx), sum(wy), sum(wz)
    gx = extent[0] * np.arange(Nx + 1) / Nx
    eta = np.arange(Ny + 1) / Ny
    gy = extent[1] * (np.exp(stretch * eta) - 1.0) / (np.exp(stretch) - 1.0)
    gz = extent[2] * np.arange(Nz + 1) / Nz
    ox = np.concatenate([[0], np.cumsum(wx)]); oy = np.concatenate([[0], np.cumsum(wy)])
    oz = np.concatenate([[0], np.cumsum(wz)])
    nbx, nby, nbz = len(wx), len(wy), len(wz)

    def bid(a, b, c):
        return a + nbx * (b + nby * c)

    dims, nodes, recs = [], [], []
    for c in range(nbz):
        for b in range(nby):
            for a in range(nbx):
                n1, n2, n3 = wx[a], wy[b], wz[c]
                dims.append((n1, n2, n3))
                xs, ys, zs = gx[ox[a]:ox[a] + n1 + 1], gy[oy[b]:oy[b] + n2 + 1], gz[oz[c]:oz[c] + n3 + 1]
                Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
                nodes.append(np.stack([X, Y, Z], axis=-1))
                ft, don = {}, {}
                nbr = {1: (a - 1, b, c), 2: (a + 1, b, c), 3: (a, b - 1, c), 4: (a, b + 1, c), 5: (a, b, c - 1),
                       6: (a, b, c + 1)}
                for f, (aa, bb, cc) in nbr.items():
                    if 0 <= aa < nbx and 0 <= bb < nby and 0 <= cc < nbz:
                        ft[f] = BC_CONNECTION
                        m1, m2, m3 = wx[aa], wy[bb], wz[cc]
                        fs = f + 1 if f % 2 == 1 else f - 1
                        if f == 1:
                            fmap = (lambda m1: lambda i, j, k: (np.full_like(i, m1), j, k))(m1)
                        elif f == 2:
                            fmap = lambda i, j, k: (np.full_like(i, 1), j, k)
                        elif f == 3:
                            fmap = (lambda m2: lambda i, j, k: (i, np.full_like(j, m2), k))(m2)
                        elif f == 4:
                            fmap = lambda i, j, k: (i, np.full_like(j, 1), k)
                        elif f == 5:
                            fmap = (lambda m3: lambda i, j, k: (i, j, np.full_like(k, m3)))(m3)
                        else:
                            fmap = lambda i, j, k: (i, j, np.full_like(k, 1))
                        don[f] = (bid(aa, bb, cc), fs, fmap, False)
                    elif f == 3:
                        ft[f] = wall
                    else:
                        ft[f] = BC_FARFIELD
                r = block_records(bid(a, b, c), (n1, n2, n3), ft, don)
                r["ff"] = np.where(r["typ"] == BC_FARFIELD, 0, -1)
                r["tw"] = np.where(r["typ"] == BC_WALL_ISOTHERMAL, T_wall, 0.0)
                recs.append(r)
    rec = np.ascontiguousarray(np.concatenate(recs))
    layout = Layout(dims)
    records_tangent(layout, nodes, rec)
    geom = metrics_build(layout, nodes, rec, ndir=3)
    # far field: stagnation state of the free stream (MOSE_open 407: T0, p0, pamb, alpha, beta, rel_fac, ci, rans)
    T0 = T_inf * (1.0 + 0.5 * (gam - 1.0) * mach ** 2)
    p0 = p_inf * (T0 / T_inf) ** (gam / (gam - 1.0))
    ff = np.zeros(1, dtype=FARFIELD_DTYPE)
    ff["mach"] = 0.0; ff["t0"] = T0; ff["p0"] = p0; ff["pamb"] = p_inf; ff["rel_fac"] = 1.0
    ff["ci"][0, 0] = 1.0
    ff["rans"][0, 0] = 3.0 * mu_inf / rho_inf
    sch = make_scheme(nrans=1, limiter=limiter, riemann=riemann, viscous=True, ndir=3, cfl=cfl, vnn=vnn,
                      prandtl=0.72, point_implicit=point_implicit)
    P = np.zeros((6, layout.NC))
    delta = 0.02 * extent[1]
    for bix, nd in enumerate(nodes):
        ctr = 0.125 * (nd[:-1, :-1, :-1] + nd[1:, :-1, :-1] + nd[:-1, 1:, :-1] + nd[:-1, :-1, 1:] + nd[1:, 1:, :-1]
                       + nd[1:, :-1, 1:] + nd[:-1, 1:, 1:] + nd[1:, 1:, 1:])          # (n3, n2, n1, 3)
        yc = ctr[..., 1]
        V = layout.interior(P, bix)
        pert = rng.uniform(-1.0, 1.0, size=(3,) + yc.shape)
        V[0] = rho_inf
        V[1] = U * np.tanh(yc / delta) + 1e-3 * U * pert[0]
        V[2] = 1e-3 * U * pert[1]
        V[3] = 1e-3 * U * pert[2]
        V[4] = p_inf
        V[5] = mu_inf * rng.uniform(1.0, 5.0, size=yc.shape)
    # the wall distance as the reference computes it (Compute_Yn: the nearest wall-face centroid; here the cell
    # centre's height above the wall up to rounding)
    wall_distance(layout, nodes, rec, geom.yn)
    fill_ghosts_nearest(layout, P)
    return Case("channel_lattice", tables, sch, layout, nodes, rec, geom, P, nsteps,
                meta={"U": U, "rho_inf": rho_inf, "mu_inf": mu_inf, "widths": (wx, wy, wz), "T0": T0, "p0": p0},
                farfield=ff)


def box_case(dims, faces, extent=(1.0, 1.0, 1.0), nrans: int = 0, viscous: bool = True, mu_of_T=None,
             prandtl: float = 0.72, farfield=None, T_wall: float = 300.0, ndir: int = 3, nsteps: int = 1,
             nodes=None, **scheme) -> Case:
    """One Cartesian block [0, extent] with a boundary type per face: faces[f] (f = 1..6) is a type id, or "periodic"
    for a connection to the opposite face (both faces of the pair must say so).  farfield: one mose_farfield row
    (FARFIELD_DTYPE) used by every type-4 face.  nodes: the block's nodes (shape (n3+1, n2+1, n1+1, 3)) in place of
    the uniform Cartesian ones.  The state is left zero (set it with layout.interior)."""
    n1, n2, n3 = dims
    nprim = 5 + nrans
    tables = air_tables(mu_of_T if mu_of_T is not None else (sutherland(1.8e-5, 300.0) if viscous else None),
                        prandtl=prandtl)
    layout = Layout([dims])
    nodes = [box_nodes(n1, n2, n3, (0.0, 0.0, 0.0), extent) if nodes is None else np.ascontiguousarray(nodes)]
    opp = {1: (2, lambda i, j, k: (np.full_like(i, n1), j, k)), 2: (1, lambda i, j, k: (np.full_like(i, 1), j, k)),
           3: (4, lambda i, j, k: (i, np.full_like(j, n2), k)), 4: (3, lambda i, j, k: (i, np.full_like(j, 1), k)),
           5: (6, lambda i, j, k: (i, j, np.full_like(k, n3))), 6: (5, lambda i, j, k: (i, j, np.full_like(k, 1)))}
    ft, don = {}, {}
    for f in range(1, 7):
        if faces[f] == "periodic":
            ft[f] = BC_CONNECTION
            fs, fmap = opp[f]
            don[f] = (0, fs, fmap, True)
        else:
            ft[f] = int(faces[f])
    rec = np.ascontiguousarray(block_records(0, dims, ft, don))
    rec["ff"] = np.where(rec["typ"] == BC_FARFIELD, 0, -1)
    rec["tw"] = np.where(rec["typ"] == BC_WALL_ISOTHERMAL, T_wall, 0.0)
    records_tangent(layout, nodes, rec)
    geom = metrics_build(layout, nodes, rec, ndir=ndir)
    kw = dict(limiter="minmod", riemann="HLLC", viscous=viscous, ndir=ndir, prandtl=prandtl if viscous else 0.0)
    kw.update(scheme)
    sch = make_scheme(nrans=nrans, **kw)
    ff = farfield if farfield is not None else np.zeros(0, dtype=FARFIELD_DTYPE)
    return Case("box", tables, sch, layout, nodes, rec, geom, np.zeros((nprim, layout.NC)), nsteps, farfield=ff)


def riemann2d(n: int = 256, nsteps: int = 100, seed: int = 0, cfl: float = 0.5, limiter="minmod",
              riemann="HLLC") -> Case:
    """2D Euler four-quadrant Riemann problem on [0, 1]^2 (one cell thick in z, extent 1/n): seeded quadrant states
    rho, p ~ U(0.1, 2), velocity of magnitude <= 1 at a seeded angle, split at (0.5, 0.5); the constant-cp air
    table, extrapolation (MOSE 400) on the four in-plane faces and no-flux faces (MOSE 0) in the one-cell direction,
    as the reference's own 2D cases mark them (test/2D/euler/isentropic-vortex/INPUT/bc.txt), SSP-RK3, two
    computational directions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tables = air_tables()
    dims = (n, n, 1)
    layout = Layout([dims])
    nodes = [box_nodes(n, n, 1, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0 / n))]
    rec = block_records(0, dims, {1: BC_EXTRAPOLATION, 2: BC_EXTRAPOLATION, 3: BC_EXTRAPOLATION,
                                  4: BC_EXTRAPOLATION, 5: BC_NONE, 6: BC_NONE})
    geom = metrics_build(layout, nodes, rec, ndir=2)
    sch = make_scheme(limiter=limiter, riemann=riemann, viscous=False, ndir=2, cfl=cfl)
    P = np.zeros((5, layout.NC))
    q = []
    for _ in range(4):
        rho, p = rng.uniform(0.1, 2.0, size=2)
        mag, ang = rng.uniform(0.0, 1.0), rng.uniform(0.0, 2 * np.pi)
        q.append((rho, mag * np.cos(ang), mag * np.sin(ang), p))
    xc = (np.arange(n) + 0.5) / n
    Y, X = np.meshgrid(xc, xc, indexing="ij")
    quad = (X >= 0.5).astype(int) + 2 * (Y >= 0.5).astype(int)
    V = layout.interior(P, 0)            # (5, 1, n, n)
    for k in range(4):
        msk = quad == k
        V[0, 0][msk], V[1, 0][msk], V[2, 0][msk], V[4, 0][msk] = q[k][0], q[k][1], q[k][2], q[k][3]
    fill_ghosts_nearest(layout, P)
    return Case("riemann2d", tables, sch, layout, nodes, rec, geom, P, nsteps, meta={"states": q, "seed": seed})


def h2_air_tables(viscous: bool = False, prandtl: float = 0.72) -> ThermoTables:
    """The flow tables of the nine hydrogen-air species of mosegpu.chem (cp, h per Kelvin on 0..10000 K, as FLINT's
    tables), and for viscous cases a power-law viscosity per species, mu_i = mu0_i (T / 300 K)^0.7 with mu0 of the
    order of the gases' 300 K values, conductivity from the Prandtl number (the transport rows are fixtures like the
    mechanism: plausible, not a validated model)."""
    from . import chem
    m = chem.h2_air_mechanism()
    mu = k = None
    if viscous:
        mu0 = np.array([8.9e-6, 2.06e-5, 1.0e-5, 9.0e-6, 1.9e-5, 1.9e-5, 2.0e-5, 1.5e-5, 1.78e-5])
        T = np.maximum(np.arange(NTAB, dtype=np.float64), 1.0)
        mu = np.ascontiguousarray(mu0[:, None] * (T[None, :] / 300.0) ** 0.7)
        k = np.ascontiguousarray(mu * m.cp / prandtl)
    return ThermoTables(m.wm, m.cp, m.h, mu, k, names=m.names)


def reacting_box(dims=(12, 10, 8), nsteps: int = 4, seed: int = 11, viscous: bool = False, phi: float = 1.0,
                 T_range=(1100.0, 1500.0), riemann="HLLC", limiter="minmod", cfl: float = 0.5) -> Case:
    """A periodic box of hydrogen-air (nine species, 1 cm on a side) with a seeded temperature field in T_range, a
    trace of H, O and OH, a swirling velocity field of 50 m/s and a 2 % pressure wave about 1 atm: the transport
    and the chemistry of a Strang-split step both act in every cell."""
    from . import chem
    tables = h2_air_tables(viscous)
    n1, n2, n3 = dims
    ext = (0.01, 0.01 * n2 / n1, 0.01 * n3 / n1)
    c = box_case(dims, {f: "periodic" for f in range(1, 7)}, extent=ext, viscous=viscous, ndir=3, nsteps=nsteps,
                 riemann=riemann, limiter=limiter, cfl=cfl, prandtl=0.72 if viscous else 0.0)
    c.tables = tables
    c.scheme.nsc = 9
    c.P = np.zeros((13, c.layout.NC))
    c.name = "reacting_box"
    rng = np.random.Generator(np.random.PCG64(seed))
    x = (np.arange(n1) + 0.5) / n1
    y = (np.arange(n2) + 0.5) / n2
    z = (np.arange(n3) + 0.5) / n3
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    a = rng.uniform(0.0, 2 * np.pi, 3)
    T = T_range[0] + (T_range[1] - T_range[0]) * 0.5 * (1 + np.sin(2 * np.pi * X + a[0]) * np.cos(2 * np.pi * Y + a[1]))
    p = 101325.0 * (1.0 + 0.02 * np.sin(2 * np.pi * Z + a[2]))
    S = chem.SPECIES
    Xm = np.zeros((9,) + X.shape)
    tot = 2 * phi + 1 + 3.76
    Xm[S.index("H2")] = 2 * phi / tot
    Xm[S.index("O2")] = 1 / tot
    Xm[S.index("N2")] = 3.76 / tot
    for sp in ("H", "O", "OH"):
        Xm[S.index(sp)] = 1e-4
    Xm /= Xm.sum(axis=0, keepdims=True)
    V = c.layout.interior(c.P, 0)
    for q in range(9):
        V[q] = Xm[q] * chem.WM[q] * p / (chem.RU * T)
    V[9] = 50.0 * np.sin(2 * np.pi * Y)
    V[10] = 30.0 * np.cos(2 * np.pi * X)
    V[11] = 10.0 * np.sin(2 * np.pi * Z)
    V[12] = p
    fill_ghosts_nearest(c.layout, c.P)
    c.meta.update({"phi": phi, "T_range": T_range})
    return c
