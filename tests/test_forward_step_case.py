"""The reference's supersonic forward step (test/2D/euler/supersonic-forward-step: two blocks joined by 101
connections, MOSE_open 405 supersonic inlets, symmetry walls, an extrapolation exit; HLLE + van Leer) read by
caseio.load_case, and its 405 inlet pinned by the oracle without the reference.

- The reader: a 405 line "mach T p alpha beta rel_fac ci" (IO_BC_Open.f90:286-300) becomes one far-field row with
  the static T and p, the supersonic marker pamb = -2.71, and the angles of the token "normal" resolved from the
  face normal (Resolve_Open_BC_Direction, IO_BC_Open.f90:411-451): 0 on the inlet faces, which are normal to x.  On a
  sheared block, whose i-faces are tilted, the resolved angle is the normal's own, atan(-1/2).
- The inlet holds its own state: with the whole field set to the row's static state at its Mach number (rho = p / (R
  T), |u| = M sqrt(gamma R T), the textbook definitions, gamma = cp / (cp - R) from the case's table), the inlet flux
  is the interior flux and no cell upstream of the step changes.  Reading T as a stagnation temperature, dropping the
  Mach factor or a wrong inflow direction each move the first column at once."""
import os

import numpy as np

from mosegpu import caseio
from mosegpu.core import BC_FARFIELD
from oracle_bind import OracleCase

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "forward-step")


def test_reader_maps_the_405_inlet():
    c = caseio.load_case(GOLD, 0)
    assert c.layout.dims == [(48, 16, 1), (240, 64, 1)]
    assert c.meta["time_threshold"] == 4.0
    inlet = c.records[c.records["typ"] == BC_FARFIELD]
    assert len(inlet) == 16 + 64 and np.all(inlet["f"] == 1) and np.all(inlet["i"] == 1)
    assert len(c.farfield) == 1 and np.all(inlet["ff"] == 0)
    row = c.farfield[0]
    assert (row["mach"], row["t0"], row["p0"], row["pamb"], row["rel_fac"]) == (3.0, 1.0, 1.0, -2.71, 1.0)
    assert row["alpha"] == 0.0 and row["beta"] == 0.0 and row["ci"][0] == 1.0


def test_inlet_holds_its_own_supersonic_state():
    c = caseio.load_case(GOLD, 0)
    row = c.farfield[0]
    R = caseio.RUNIV / c.tables.wm[0]
    T, p, M = row["t0"], row["p0"], row["mach"]
    it = int(T)
    cp = c.tables.cp[0, it] + (T - it) * (c.tables.cp[0, it + 1] - c.tables.cp[0, it])
    gam = cp / (cp - R)
    rho, u = p / (R * T), M * np.sqrt(gam * R * T)
    for b in range(c.layout.nb):
        V = c.layout.interior(c.P, b)
        V[0], V[1], V[2], V[3], V[4] = rho, u, 0.0, 0.0, p
    caseio.fill_ghosts_nearest(c.layout, c.P)
    P = c.P.copy()
    OracleCase(c, nthreads=8).step(P, 5)
    scale = np.array([rho, u, u, u, p])[:, None, None, None]
    for b in range(c.layout.nb):
        d = np.abs(c.layout.interior(P, b) - c.layout.interior(c.P, b)) / scale
        assert np.max(d[..., :30]) <= 1e-14, (b, np.max(d[..., :30]))
    # the step is there: the cells in front of it do change
    d0 = np.abs(c.layout.interior(P, 0) - c.layout.interior(c.P, 0)) / scale
    assert np.max(d0[..., -2:]) > 1e-3


def test_inlet_angles_from_the_face_normal():
    """On a sheared block (x = X + Y / 2) the i-faces' normal is (1, -1/2, 0) / |.|: "normal" tokens give alpha =
    atan(-1/2), beta = 0; a number is taken as it is (a Fortran d exponent included); equal lines share one row."""
    from mosegpu import cases

    n = (4, 3, 2)
    nodes = cases.box_nodes(*n, (0.0, 0.0, 0.0), (1.0, 0.9, 0.4))
    nodes[..., 0] += 0.5 * nodes[..., 1]
    c = cases.box_case(n, {f: BC_FARFIELD for f in range(1, 7)}, viscous=False, nodes=nodes)
    rec = np.ascontiguousarray(c.records.copy())
    face1 = np.flatnonzero(rec["f"] == 1)
    face3 = np.flatnonzero(rec["f"] == 3)[0]
    inl = [(int(e), ["3.0,", "1.0,", "1.0,", "normal,", "normal,", "1.0,", "1.0,"]) for e in face1]
    inl.append((int(face3), ["2.0", "1.0", "1.0", "0.25d0", "normal", "1.0", "1.0"]))
    ff = caseio.inlet_rows(rec, inl, c.geom, c.layout, 1, 0)
    assert len(ff) == 2 and np.all(ff["pamb"] == -2.71)
    assert np.all(rec["ff"][face1] == 0) and rec["ff"][face3] == 1
    assert abs(ff[0]["alpha"] - np.arctan(-0.5)) <= 1e-15 and abs(ff[0]["beta"]) <= 1e-15
    assert (ff[1]["mach"], ff[1]["alpha"], ff[1]["beta"]) == (2.0, 0.25, 0.0)
