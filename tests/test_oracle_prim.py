"""The oracle's primitive-variable update (integration-variables = prim, Newstate_Primitive_blk,
src/lib/numerics/time/explicit/Lib_Newstate.f90:340-531), pinned without the reference.

Newstate_Primitive_blk turns the conservative residual into primitive increments through the inverse of
d(cons)/d(prim) at the current state.  For a forward-Euler stage (PO = P, rkc = 1) it is therefore the linearisation
of the conservative update: with the residual scaled by eps, the two updates differ by O(eps^2).  On calorically
perfect tables (cp constant, h linear in T; the conservative update's Newton temperature is then exact) the ratio of
the velocity and pressure differences at eps and eps/2 is 4.  Any wrong entry of the inverse (a sign, a dropped term,
Theta, the mass-fraction sensitivities of the several-species branch) leaves an O(eps) difference, a ratio of 2.  The
species densities are linear in both sets of variables (rho_p dp + rho_T dT + sum drho/dY_s dY_s rebuilds the
conservative mass increment), so there the two updates agree to rounding, and so does the Spalart-Allmaras variable,
which both advance by the conservative increment."""
import numpy as np
import pytest

from mosegpu import cases
from mosegpu.core import ThermoTables, NTAB
from oracle_bind import OracleCase


def _box(nsc, nrans, integration, tables=None, seed=5):
    c = cases.box_case((5, 4, 3), {f: "periodic" for f in range(1, 7)}, viscous=False, nrans=nrans,
                       integration=integration)
    if nsc == 1 and tables is None:
        return c
    sch = cases.make_scheme(nsc=nsc, nrans=nrans, viscous=False, integration=integration)
    nprim = nsc + 4 + nrans
    P = np.zeros((nprim, c.layout.NC))
    return cases.Case("box", tables, sch, c.layout, c.nodes, c.records, c.geom, P, c.nsteps)


def _perfect_species():
    """Three calorically perfect species with different molar masses, heat capacities and formation enthalpies."""
    T = np.arange(NTAB, dtype=np.float64)
    wm = np.array([28.0, 32.0, 18.0])
    cp = np.array([1040.0, 918.0, 1860.0])
    h0 = np.array([0.0, 3.0e4, -1.3e7])
    return ThermoTables(wm, np.repeat(cp[:, None], NTAB, axis=1), h0[:, None] + cp[:, None] * T[None, :])


def _state(case, nsc, nrans, rng):
    V = case.layout.interior(case.P, 0)
    sh = V.shape[1:]
    rho = 1.2 * (1 + 0.1 * rng.uniform(-1, 1, sh))
    Y = rng.uniform(0.2, 1.0, (nsc,) + sh)
    Y /= Y.sum(axis=0)
    for s in range(nsc):
        V[s] = rho * Y[s]
    for q in range(3):
        V[nsc + q] = 40.0 * rng.uniform(-1, 1, sh)
    V[nsc + 3] = 1.0e5 * (1 + 0.1 * rng.uniform(-1, 1, sh))
    if nrans:
        V[nsc + 4] = 1e-4 * rho * (1 + 0.5 * rng.uniform(0, 1, sh))


def _updates(nsc, nrans, tables, eps_list, seed=5):
    """P after one forward-Euler update with the residual eps * R0, primitive and conservative, per eps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cp_ = _box(nsc, nrans, "prim", tables)
    cc_ = _box(nsc, nrans, "cons", tables)
    _state(cp_, nsc, nrans, rng)
    cc_.P[:] = cp_.P
    nprim = nsc + 4 + nrans
    R0 = np.zeros_like(cp_.P)
    V = cp_.layout.interior(cp_.P, 0)
    vol = cp_.layout.interior(cp_.geom.vol[None, :], 0)[0]
    # a residual of the size of the state: eps = 1 changes every conserved variable by O(1) of itself
    rho = V[:nsc].sum(axis=0)
    E = V[nsc + 3] / 0.4 + 0.5 * rho * (V[nsc] ** 2 + V[nsc + 1] ** 2 + V[nsc + 2] ** 2)
    scale = [V[s] for s in range(nsc)] + [rho * 40.0] * 3 + [E] + ([V[nsc + 4]] if nrans else [])
    RI = cp_.layout.interior(R0, 0)
    for s in range(nprim):
        RI[s] = scale[s] * vol * rng.uniform(-1, 1, V.shape[1:])
    op, oc = OracleCase(cp_, nthreads=1), OracleCase(cc_, nthreads=1)
    out = []
    for eps in eps_list:
        Pp, Pc = cp_.P.copy(), cc_.P.copy()
        op.newstate(Pp, cp_.P, eps * R0, 1.0, 1.0)
        oc.newstate(Pc, cc_.P, eps * R0, 1.0, 1.0)
        out.append((cp_.layout.interior(Pp, 0).copy(), cp_.layout.interior(Pc, 0).copy()))
    return V.copy(), out


def _ratios(V, out):
    """per variable: max |prim - cons| / max |P| at eps, and its ratio to the same at eps / 2"""
    nprim = V.shape[0]
    d = [[np.max(np.abs(p[s] - c[s])) / np.max(np.abs(V[s])) for s in range(nprim)] for p, c in out]
    return d, [[d[e][s] / d[e + 1][s] if d[e + 1][s] > 0 else np.inf for s in range(nprim)] for e in range(len(d) - 1)]


@pytest.mark.parametrize("nsc", [1, 3])
def test_prim_update_is_the_linearised_conservative_update(nsc):
    tables = cases.air_tables() if nsc == 1 else _perfect_species()
    V, out = _updates(nsc, 0, tables, [4e-3, 2e-3, 1e-3])
    d, r = _ratios(V, out)
    for s in range(nsc):                                # species densities: the same increment, to rounding
        assert max(de[s] for de in d) <= 2e-15, (nsc, s, d)
    for s in range(nsc, V.shape[0]):                    # velocity and pressure: second order in eps
        assert d[0][s] > 1e-9, (s, d[0][s])             # far above rounding: the ratio measures the order
        for e in range(len(r)):
            assert 3.6 < r[e][s] < 4.4, (nsc, s, d, r)


def test_prim_update_turbulence_variable_is_the_conservative_increment():
    """nu~ of a Spalart-Allmaras box: the primitive update adds -R dt / vol, as the conservative one does."""
    V, out = _updates(1, 1, cases.air_tables(), [4e-3])
    p, c = out[0]
    nt = 5
    assert np.max(np.abs(p[nt] - c[nt])) <= 4e-15 * np.max(np.abs(V[nt]))
    assert np.max(np.abs(p[nt] - V[nt])) > 1e-4 * np.max(np.abs(V[nt]))     # and the increment is not zero


def test_prim_update_of_a_zero_residual_keeps_the_state():
    for nsc, tab in ((1, cases.air_tables()), (3, _perfect_species())):
        V, out = _updates(nsc, 0, tab, [0.0])
        assert np.array_equal(out[0][0], V) and np.array_equal(out[0][1][:nsc], V[:nsc])
