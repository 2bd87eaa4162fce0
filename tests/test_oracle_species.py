"""Pins of the oracle's multi-species residual (four species) on the CPU.

- Four copies of one gas, in fixed mass fractions, are that gas.  Each mixture sum the residual takes then reduces to
  the one-species value: rho = sum rho_i, R, cp, h, and Wilke's mixing rule with phi_ij = 1.  The runs agree to
  rounding on a viscous periodic box in motion.
- Binary diffusion at rest.  Two copies of one gas carry a sine mode of mass fraction in a periodic 1D box, at
  uniform p and T.  The mode decays as exp(-D k_h^2 t), with D = mu / (Sc rho) and k_h^2 = (2 - 2 cos kh) / h^2 (the
  three-point Laplacian of the face gradients).  The other species stay uniform and the gas stays at rest.
- The reference's own test/1D/diff1 inputs (CH4, H2O, N2, O2 layers at rest, periodic and symmetry faces): each
  species' mass is conserved to rounding and the pressure stays within 1e-4 of its mean.
"""
import os

import numpy as np
import pytest

from mosegpu import caseio, cases
from mosegpu.core import BC_NONE, Layout, ThermoTables, metrics_build
from oracle_bind import OracleCase

HERE = os.path.dirname(os.path.abspath(__file__))
DIFF1 = os.path.join(HERE, "golden", "cases", "diff1")
Y4 = np.array([0.1, 0.2, 0.3, 0.4])


def copies(t: ThermoTables, n: int) -> ThermoTables:
    rep = lambda a: None if a is None else np.repeat(a, n, axis=0)
    return ThermoTables(np.repeat(t.wm, n), rep(t.cp), rep(t.h), rep(t.mu), rep(t.k))


def as_four_species(c1: cases.Case, Y, rs: str) -> cases.Case:
    sch = cases.make_scheme(nsc=4, limiter="vanleer", riemann=rs, viscous=bool(c1.scheme.viscous), ndir=c1.scheme.ndir,
                            prandtl=c1.scheme.prandtl, sc=c1.scheme.sc)
    P = np.zeros((8, c1.layout.NC))
    for s in range(4):
        P[s] = Y[s] * c1.P[0]
    P[4:] = c1.P[1:5]
    return cases.Case("four", copies(c1.tables, 4), sch, c1.layout, c1.nodes, c1.records, c1.geom, P, c1.nsteps)


@pytest.mark.parametrize("rs", ["HLLC", "LMRoe"])
def test_four_copies_are_one_gas(rs):
    c1 = cases.box_case((6, 5, 4), {f: "periodic" for f in range(1, 7)}, extent=(1.0, 0.9, 0.8), viscous=True,
                        mu_of_T=cases.sutherland(1.8e-3, 300.0), riemann=rs, limiter="vanleer")
    rng = np.random.Generator(np.random.PCG64(3))
    V = c1.layout.interior(c1.P, 0)
    V[0] = 1.1 + 0.1 * rng.uniform(-1, 1, V[0].shape)
    for q in range(3):
        V[1 + q] = 30.0 * rng.uniform(-1, 1, V[0].shape)
    V[4] = 1e5 * (1.0 + 0.05 * rng.uniform(-1, 1, V[0].shape))
    cases.fill_ghosts_nearest(c1.layout, c1.P)
    c4 = as_four_species(c1, Y4, rs)
    P1, P4 = c1.P.copy(), c4.P.copy()
    d1 = OracleCase(c1).step(P1, 4)
    d4 = OracleCase(c4).step(P4, 4)
    assert abs(d1.sum() - d4.sum()) <= 1e-12 * d1.sum()
    A, B = c1.layout.interior(P1, 0), c4.layout.interior(P4, 0)
    rho4 = B[0] + B[1] + B[2] + B[3]
    for a, b in [(A[0], rho4)] + [(A[q], B[3 + q]) for q in range(1, 5)]:
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
    for s in range(4):      # the mass fractions stay what they were
        assert np.max(np.abs(B[s] / rho4 - Y4[s])) <= 1e-13


def test_binary_diffusion_mode_decay():
    n, L, p, T, Sc, amp = 48, 1.0, 1.0e5, 300.0, 0.7, 0.1
    air = cases.air_tables(cases.sutherland(50.0, 300.0), prandtl=0.72)   # diffusion fast against sound
    tab = copies(air, 4)
    layout = Layout([(n, 1, 1)])
    nodes = [cases.box_nodes(n, 1, 1, (0.0, 0.0, 0.0), (L, 0.1, 0.1))]
    faces = {1: cases.BC_CONNECTION, 2: cases.BC_CONNECTION, 3: BC_NONE, 4: BC_NONE, 5: BC_NONE, 6: BC_NONE}
    donors = {1: (0, 2, lambda i, j, k: (np.full_like(i, n), j, k), True),
              2: (0, 1, lambda i, j, k: (np.full_like(i, 1), j, k), True)}
    rec = np.ascontiguousarray(cases.block_records(0, (n, 1, 1), faces, donors))
    geom = metrics_build(layout, nodes, rec, ndir=1)
    sch = cases.make_scheme(nsc=4, limiter="vanleer", riemann="HLLC", viscous=True, ndir=1, prandtl=0.72, sc=Sc,
                            cfl=0.5, vnn=0.3)
    x = (np.arange(n) + 0.5) * L / n
    rho = p / (cases.R_AIR * T)
    k = 2.0 * np.pi / L
    Y = [0.3 + amp * np.sin(k * x), 0.3 - amp * np.sin(k * x), np.full(n, 0.2), np.full(n, 0.2)]
    P = np.zeros((8, layout.NC))
    V = layout.interior(P, 0)
    for s in range(4):
        V[s] = rho * Y[s]
    V[7] = p
    cases.fill_ghosts_nearest(layout, P)
    case = cases.Case("diffusion", tab, sch, layout, nodes, rec, geom, P, 0)
    Pn = P.copy()
    t = OracleCase(case).step(Pn, 60).sum()
    W = layout.interior(Pn, 0)[:, 0, 0, :]
    rt = W[0] + W[1] + W[2] + W[3]
    a = 2.0 / n * np.sum(W[0] / rt * np.sin(k * x))
    h = L / n
    D = air.mu[0, 300] / (Sc * rho)
    want = amp * np.exp(-D * (2.0 - 2.0 * np.cos(k * h)) / h ** 2 * t)
    assert want < 0.8 * amp                                          # the mode has visibly decayed
    assert abs(a - want) <= 1e-6 * amp, (a, want)
    assert np.max(np.abs(W[2] / rt - 0.2)) <= 1e-13 and np.max(np.abs(W[4:7])) <= 1e-9
    assert np.max(np.abs(W[7] - p)) <= 1e-9 * p


def test_diff1_conserves_species():
    case = caseio.load_case(DIFF1, 100)
    assert case.scheme.nsc == 4 and 3 in set(case.records["typ"].tolist())
    vol = case.layout.interior(case.geom.vol[None, :], 0)[0]
    V0 = case.layout.interior(case.P, 0)
    m0 = np.array([np.sum(V0[s] * vol) for s in range(4)])
    P = case.P.copy()
    OracleCase(case).step(P, 100)
    V = case.layout.interior(P, 0)
    m = np.array([np.sum(V[s] * vol) for s in range(4)])
    assert np.all(np.abs(m - m0) <= 1e-14 * m0.max())
    assert np.max(np.abs(V[7] - V[7].mean())) <= 1e-4 * V[7].mean()
