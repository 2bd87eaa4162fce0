"""Exact solution of the 1D Riemann problem of an ideal gas (Toro, "Riemann Solvers and Numerical Methods for
Fluid Dynamics", 3rd ed., ch. 4: pressure function, Newton iteration on p*, sampling of the wave pattern).
Independent of the reference; used to pin the oracle on the Sod problem."""
from __future__ import annotations

import numpy as np


def _f(p, rk, pk, ck, g):
    if p > pk:   # shock
        A = 2.0 / ((g + 1.0) * rk)
        B = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    pr = p / pk   # rarefaction
    return 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0), pr ** (-(g + 1.0) / (2.0 * g)) / (rk * ck)


def star_state(rL, uL, pL, rR, uR, pR, g=1.4):
    cL, cR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
    p = max(1e-12, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (cL + cR))
    for _ in range(100):
        fL, dL = _f(p, rL, pL, cL, g)
        fR, dR = _f(p, rR, pR, cR, g)
        dp = (fL + fR + uR - uL) / (dL + dR)
        pn = max(1e-12, p - dp)
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fL, _ = _f(p, rL, pL, cL, g)
    fR, _ = _f(p, rR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(x, t, x0, rL, uL, pL, rR, uR, pR, g=1.4):
    """(rho, u, p) at the points x at time t for the discontinuity at x0."""
    ps, us = star_state(rL, uL, pL, rR, uR, pR, g)
    cL, cR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
    out = np.zeros((3, len(x)))
    for n, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:   # left of the contact
            if ps > pL:
                SL = uL - cL * np.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                if s <= SL:
                    out[:, n] = (rL, uL, pL)
                else:
                    r = rL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1))
                    out[:, n] = (r, us, ps)
            else:
                cs = cL * (ps / pL) ** ((g - 1) / (2 * g))
                if s <= uL - cL:
                    out[:, n] = (rL, uL, pL)
                elif s >= us - cs:
                    out[:, n] = (rL * (ps / pL) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (cL + (g - 1) / 2 * uL + s)
                    c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - s))
                    r = rL * (c / cL) ** (2 / (g - 1))
                    out[:, n] = (r, u, pL * (c / cL) ** (2 * g / (g - 1)))
        else:
            if ps > pR:
                SR = uR + cR * np.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                if s >= SR:
                    out[:, n] = (rR, uR, pR)
                else:
                    r = rR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
                    out[:, n] = (r, us, ps)
            else:
                cs = cR * (ps / pR) ** ((g - 1) / (2 * g))
                if s >= uR + cR:
                    out[:, n] = (rR, uR, pR)
                elif s <= us + cs:
                    out[:, n] = (rR * (ps / pR) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + s)
                    c = 2 / (g + 1) * (cR - (g - 1) / 2 * (uR - s))
                    r = rR * (c / cR) ** (2 / (g - 1))
                    out[:, n] = (r, u, pR * (c / cR) ** (2 * g / (g - 1)))
    return out
