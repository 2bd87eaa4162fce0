"""A 9-species H2-O2-N2 finite-rate mechanism in FLINT's table form, and seeded batches of reacting cells.

The reference has no 9-species hydrogen data files: FLINT's table-driven `general` path reads chemistry-info.txt /
chemistry-Arrhenius.dat / chemistry-Troe.dat (Load_Chemistry.f90:7-164), and none ships for this set. This module
authors them. It holds 19 reactions in the form of Li et al. (Int. J. Chem. Kinet. 36, 2004): 17 Arrhenius, of which
4 carry a third body, and 2 Troe falloff reactions.

The thermodynamics are smooth fits, per species: cp/R = a0 + a1 t + a2 t^2 with t = T/1000 clamped to [0.2, 4],
plus the standard formation enthalpies and entropies at 298.15 K.

The reverse rates come from the same tables by detailed balance: kb = kf / Kc, with
Kc = exp(-dG0 / (Ru T)) (p0 / (Ru T))^dn and p0 = 1e5 Pa. Every table has one row per Kelvin on 0..10000 K,
as FLINT's tables do. Rows below 200 K repeat the 200 K values.

The data are fixtures: they are physically plausible, not a validated kinetic model. Parity of the device path is
defined against the oracle on these same tables.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

RU = 8314.46261815324          # J / (kmol K)
NTAB = 10001
SPECIES = ["H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "N2"]
WM = np.array([2.01588, 31.9988, 18.01528, 1.00794, 15.9994, 17.00734, 33.00674, 34.01468, 28.0134])
# cp/R = a0 + a1 t + a2 t^2, t = T / 1000 clamped to [0.2, 4]
CPFIT = np.array([
    [3.30, 0.30, 0.00],     # H2
    [3.40, 0.70, -0.15],    # O2
    [3.80, 0.90, -0.05],    # H2O
    [2.50, 0.00, 0.00],     # H
    [2.60, -0.05, 0.00],    # O
    [3.40, 0.30, 0.00],     # OH
    [4.00, 1.00, -0.15],    # HO2
    [4.80, 2.00, -0.40],    # H2O2
    [3.35, 0.55, -0.12],    # N2
])
HF298 = np.array([0.0, 0.0, -241.826e6, 217.999e6, 249.175e6, 37.30e6, 12.02e6, -135.88e6, 0.0])       # J / kmol
S298 = np.array([130.68e3, 205.15e3, 188.84e3, 114.72e3, 161.06e3, 183.74e3, 229.0e3, 232.99e3, 191.61e3])  # J/(kmol K)
P0 = 1.0e5
CAL = 4.184                    # J / cal

# reactions: (reactants, products, A [cm, mol, s], b, Ea [cal/mol], third-body efficiencies or None)
_ARRH = [
    ({"H": 1, "O2": 1}, {"O": 1, "OH": 1}, 3.547e15, -0.406, 16599.0, None),
    ({"O": 1, "H2": 1}, {"H": 1, "OH": 1}, 5.08e4, 2.67, 6290.0, None),
    ({"H2": 1, "OH": 1}, {"H2O": 1, "H": 1}, 2.16e8, 1.51, 3430.0, None),
    ({"O": 1, "H2O": 1}, {"OH": 2}, 2.97e6, 2.02, 13400.0, None),
    ({"H2": 1}, {"H": 2}, 4.577e19, -1.40, 104380.0, {"H2": 2.5, "H2O": 12.0}),
    ({"O": 2}, {"O2": 1}, 6.165e15, -0.50, 0.0, {"H2": 2.5, "H2O": 12.0}),
    ({"O": 1, "H": 1}, {"OH": 1}, 4.714e18, -1.0, 0.0, {"H2": 2.5, "H2O": 12.0}),
    ({"H": 1, "OH": 1}, {"H2O": 1}, 3.8e22, -2.0, 0.0, {"H2": 2.5, "H2O": 12.0}),
    ({"HO2": 1, "H": 1}, {"H2": 1, "O2": 1}, 1.66e13, 0.0, 823.0, None),
    ({"HO2": 1, "H": 1}, {"OH": 2}, 7.079e13, 0.0, 295.0, None),
    ({"HO2": 1, "O": 1}, {"O2": 1, "OH": 1}, 3.25e13, 0.0, 0.0, None),
    ({"HO2": 1, "OH": 1}, {"H2O": 1, "O2": 1}, 2.89e13, 0.0, -497.0, None),
    ({"HO2": 2}, {"H2O2": 1, "O2": 1}, 4.2e14, 0.0, 11982.0, None),
    ({"H2O2": 1, "H": 1}, {"H2O": 1, "OH": 1}, 2.41e13, 0.0, 3970.0, None),
    ({"H2O2": 1, "H": 1}, {"HO2": 1, "H2": 1}, 4.82e13, 0.0, 7950.0, None),
    ({"H2O2": 1, "O": 1}, {"OH": 1, "HO2": 1}, 9.55e6, 2.0, 3970.0, None),
    ({"H2O2": 1, "OH": 1}, {"HO2": 1, "H2O": 1}, 1.0e12, 0.0, 0.0, None),
]
# Troe: (reactants, products, (A, b, Ea) high, (A, b, Ea) low, (alpha, T3, T1, T2 or None), efficiencies)
_TROE = [
    ({"H": 1, "O2": 1}, {"HO2": 1}, (1.475e12, 0.60, 0.0), (6.366e20, -1.72, 524.8), (0.8, 1e-30, 1e30, None),
     {"H2": 2.0, "H2O": 11.0, "O2": 0.78}),
    ({"H2O2": 1}, {"OH": 2}, (2.951e14, 0.0, 48430.0), (1.202e17, 0.0, 45500.0), (0.5, 1e-30, 1e30, None),
     {"H2": 2.5, "H2O": 12.0}),
]


def _t_eval():
    return np.maximum(np.arange(NTAB, dtype=np.float64), 200.0)


def thermo_molar():
    """cp [J/(kmol K)], h [J/kmol], s [J/(kmol K)] per species on the rows T = 0..10000 K (T < 200 as 200)."""
    T = _t_eval()
    t = np.clip(T / 1000.0, 0.2, 4.0)
    cp = RU * (CPFIT[:, :1] + CPFIT[:, 1:2] * t + CPFIT[:, 2:3] * t * t)
    # h and s from the 298.15 K values by the trapezoid rule on a 1 K grid (exact integrals of the piecewise-linear cp)
    Tg = np.arange(NTAB, dtype=np.float64)
    Tg = np.maximum(Tg, 200.0)
    h = np.zeros_like(cp)
    s = np.zeros_like(cp)
    for q in range(len(SPECIES)):
        c = cp[q]
        dh = np.concatenate([[0.0], np.cumsum(0.5 * (c[1:] + c[:-1]) * np.diff(Tg))])
        ds = np.concatenate([[0.0], np.cumsum(0.5 * (c[1:] / Tg[1:] + c[:-1] / Tg[:-1]) * np.diff(Tg))])
        h298 = np.interp(298.15, Tg, dh)
        s298 = np.interp(298.15, Tg, ds)
        h[q] = HF298[q] + dh - h298
        s[q] = S298[q] + ds - s298
    return cp, h, s


def _kc(react, prod, h, s):
    T = _t_eval()
    dg = np.zeros(NTAB)
    dn = 0
    for sp, nu in prod.items():
        q = SPECIES.index(sp)
        dg += nu * (h[q] - T * s[q])
        dn += nu
    for sp, nu in react.items():
        q = SPECIES.index(sp)
        dg -= nu * (h[q] - T * s[q])
        dn -= nu
    return np.exp(-dg / (RU * T)) * (P0 / (RU * T)) ** dn


def _arr(A, b, Ea, order):
    T = _t_eval()
    return A * (1e-3) ** (order - 1) * T ** b * np.exp(-Ea * CAL / (RU * T))


@dataclass
class Mechanism:
    ns: int
    wm: np.ndarray
    cp: np.ndarray      # [ns, NTAB] J/(kg K)
    h: np.ndarray       # [ns, NTAB] J/kg
    ni1: np.ndarray     # [nr, ns+1]
    ni2: np.ndarray
    eps: np.ndarray
    kf: np.ndarray      # [n_arrh, NTAB]
    kb: np.ndarray
    troe: np.ndarray    # [n_troe, 4, NTAB]
    lind: np.ndarray    # [n_lind, 3, NTAB]
    names: list

    @property
    def n_arrh(self):
        return self.kf.shape[0]

    @property
    def n_troe(self):
        return self.troe.shape[0]

    @property
    def n_lind(self):
        return self.lind.shape[0]

    def as_struct(self):
        from .core import MechanismStruct, dptr
        m = MechanismStruct()
        m.ns = self.ns; m.n_arrh = self.n_arrh; m.n_troe = self.n_troe; m.n_lind = self.n_lind
        m.tmin = 0; m.tmax = NTAB - 1
        for name in ("wm", "cp", "h", "ni1", "ni2", "eps", "kf", "kb", "troe", "lind"):
            setattr(m, name, dptr(getattr(self, name)))
        m._keep = self
        return m


def h2_air_mechanism() -> Mechanism:
    ns = len(SPECIES)
    cpm, hm, sm = thermo_molar()
    rows = []
    kf, kb, troe = [], [], []
    for react, prod, A, b, Ea, eff in _ARRH:
        order = sum(react.values()) + (1 if eff is not None else 0)
        kfr = _arr(A, b, Ea, order)
        kf.append(kfr)
        kb.append(kfr / _kc(react, prod, hm, sm))
        rows.append((react, prod, eff))
    for react, prod, hi, lo, (al, T3, T1, T2), eff in _TROE:
        order = sum(react.values())
        kinf = _arr(*hi, order)
        k0 = _arr(*lo, order + 1)
        T = _t_eval()
        fc = (1 - al) * np.exp(-T / T3) + al * np.exp(-T / T1) + (np.exp(-T2 / T) if T2 else 0.0)
        troe.append(np.stack([kinf, k0, _kc(react, prod, hm, sm), fc]))
        rows.append((react, prod, eff))
    nr = len(rows)
    ni1 = np.zeros((nr, ns + 1)); ni2 = np.zeros((nr, ns + 1)); eps = np.zeros((nr, ns + 1))
    for r, (react, prod, eff) in enumerate(rows):
        for sp, nu in react.items():
            ni1[r, SPECIES.index(sp)] = nu
        for sp, nu in prod.items():
            ni2[r, SPECIES.index(sp)] = nu
        if eff is not None:
            ni1[r, ns] = 1.0
            eps[r, :ns] = 1.0
            for sp, e in eff.items():
                eps[r, SPECIES.index(sp)] = e
    return Mechanism(ns, WM.copy(), np.ascontiguousarray(cpm / WM[:, None]), np.ascontiguousarray(hm / WM[:, None]),
                     ni1, ni2, eps, np.ascontiguousarray(np.array(kf)), np.ascontiguousarray(np.array(kb)),
                     np.ascontiguousarray(np.array(troe)), np.zeros((0, 3, NTAB)), list(SPECIES))


def h2_air_cells(n: int, seed: int = 4, T_range=(900.0, 2500.0), phi_range=(0.5, 2.0), p: float = 101325.0,
                 mech: Mechanism | None = None) -> np.ndarray:
    """n seeded cells of H2-air (O2:N2 = 1:3.76 molar): T ~ U(T_range), phi ~ U(phi_range), pressure p.
    Returns the state [ns + 1, n]: species densities (kg/m^3), then pressure."""
    rng = np.random.Generator(np.random.PCG64(seed))
    T = rng.uniform(*T_range, size=n)
    phi = rng.uniform(*phi_range, size=n)
    ns = len(SPECIES)
    X = np.zeros((ns, n))
    tot = 2 * phi + 1 + 3.76
    X[SPECIES.index("H2")] = 2 * phi / tot
    X[SPECIES.index("O2")] = 1 / tot
    X[SPECIES.index("N2")] = 3.76 / tot
    Wmix = (X * WM[:, None]).sum(axis=0)
    rho = p * Wmix / (RU * T)
    out = np.zeros((ns + 1, n))
    out[:ns] = rho * X * WM[:, None] / Wmix
    out[ns] = p
    return out


def ode_opts(rtol: float = 1e-8, atol: float = 1e-12, ns: int = 9, method: int = 1):
    """method 1: RODAS, 2: RADAU5 (MOSE_ODE_RODAS / MOSE_ODE_RADAU5); raw tolerances for every component"""
    from .core import OdeOpts
    o = OdeOpts()
    o.method = method
    for q in range(ns + 1):
        o.rtol[q] = rtol
        o.atol[q] = atol
    return o
