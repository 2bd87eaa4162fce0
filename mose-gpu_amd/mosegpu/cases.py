"""Synthetic cases of the residual path: thermo tables, meshes, boundary records and initial fields.

The cases follow the reference's case format semantics (test/1D/Sod79/INPUT: a constant-cp air table with
h = cp T, FLINT's T = 0..10000 K rows; MOSE boundary types 400 -> 11, 201 -> 1 periodic), built in memory:
  sod(n)            1D shock tube, n x 1 x 1 on x in [0, 1], minmod + HLLC, SSP-RK3, CFL 0.5, transmissive ends
  taylor_green(n)   3D periodic n^3 on [0, 2 pi]^3, M = 0.1, Re = 1600, Roe (LMRoe with Mco = 1) + van Albada,
                    Sutherland viscosity tabulated per Kelvin, Pr = 0.72, SSP-RK3, seeded velocity perturbation
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .core import (BC_CONNECTION, BC_EXTRAPOLATION, GC, LIM, NTAB, RECORD_DTYPE, RS, Geometry, Layout, Scheme,
                   ThermoTables, metrics_build)

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

    @property
    def nprim(self):
        return self.scheme.nsc + 4 + self.scheme.nrans

    def interior_cells(self):
        return sum(n1 * n2 * n3 for n1, n2, n3 in self.layout.dims)


def make_scheme(nsc=1, nrans=0, limiter="minmod", riemann="HLLC", viscous=False, n_rk=3, ndir=3, cfl=0.5,
                vnn=0.5, mco=1.0, prandtl=0.0, sc=0.5, sct=0.9, prt=0.9, point_implicit=1, guard=1) -> Scheme:
    s = Scheme()
    s.nsc = nsc; s.nrans = nrans; s.limiter = LIM[limiter]; s.riemann = RS[riemann]; s.viscous = int(viscous)
    s.n_rk = n_rk; s.ndir = ndir; s.point_implicit = point_implicit; s.thermo_guard = guard
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
                 p0: float = 101325.0, amp: float = 1e-3, seed: int = 2024, cfl: float = 0.5, vnn: float = 0.5,
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
