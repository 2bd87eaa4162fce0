"""Pins of the oracle's point functions (CPU): limiters against their closed forms on a brute-force grid,
Riemann fluxes against consistency, upwinding, the stationary contact, mirror symmetry and an independent
textbook Roe flux, and the MUSCL face against constant and linear fields."""
import numpy as np
import pytest

from mosegpu import cases
from oracle_bind import OracleCase, oracle_lib


@pytest.fixture(scope="module")
def sod_oracle():
    return OracleCase(cases.sod(40, 1))


# --- limiters: Lib_Limiters.f90:65-97 -----------------------------------------------------------------------
def _lim_ref(lim, x, y):
    if lim == 1:
        if x * y < 0.0:
            return 0.0
        return y if abs(x) > abs(y) else x
    if lim == 6:
        return 0.0
    ru = y / (x + 1e-30)
    if lim == 2:
        phi = (ru + abs(ru)) / (1.0 + abs(ru))
    elif lim == 3:
        phi = (ru * ru + ru) / (1.0 + ru * ru)
    elif lim == 4:
        phi = max(0.0, min(2.0 * ru, 0.5 * (1.0 + ru), 2.0))
    else:
        phi = max(0.0, min(2.0 * ru, 1.0), min(ru, 2.0))
    return phi * x


VALS = [0.0, -0.0, 1e-300, -1e-300, 1e-30, 3e-17, 0.5, -0.5, 1.0, -1.0, 2.0, -3.0, 1e10, -7.25, 0.999, 1.001]


@pytest.mark.parametrize("lim", [1, 2, 3, 4, 5, 6])
def test_limiter_bruteforce(lim):
    L = oracle_lib()
    for x in VALS:
        for y in VALS:
            got = L.oa_limiter(lim, x, y)
            ref = _lim_ref(lim, x, y)
            assert got == ref or (np.isnan(got) and np.isnan(ref)), (lim, x, y, got, ref)


def test_limiter_special_cases():
    L = oracle_lib()
    assert L.oa_limiter(1, 1.0, -2.0) == 0.0            # opposite signs -> 0
    assert L.oa_limiter(1, 2.0, 2.0) == 2.0             # tie -> x
    assert L.oa_limiter(1, -3.0, -2.0) == -2.0          # |x| > |y| -> y
    assert L.oa_limiter(3, 2.0, 2.0) == 2.0             # van Albada phi(1) = 1
    assert L.oa_limiter(2, 1.0, 3.0) == 1.5             # van Leer phi(3) = 6/4
    # symmetric limiters: phi(r)/r = phi(1/r) -> lim(x, y) == lim(y, x) for van Leer / van Albada / minmod
    for lim in (1, 2, 3):
        for x, y in [(0.3, 1.7), (2.0, 0.5), (-1.0, -4.0)]:
            assert abs(L.oa_limiter(lim, x, y) - L.oa_limiter(lim, y, x)) <= 4e-16 * abs(x + y)


# --- Riemann fluxes -----------------------------------------------------------------------------------------
def phys_flux(q, n, cp=1004.5, R=8314.46261815324 / 28.970418):
    r, u, v, w, p = q
    un = u * n[0] + v * n[1] + w * n[2]
    h = cp * (p / (r * R))
    H = h + 0.5 * (u * u + v * v + w * w)
    return np.array([r * un, r * un * u + p * n[0], r * un * v + p * n[1], r * un * w + p * n[2], r * un * H])


STATES = [np.array([1.0, 0.3, -0.2, 0.1, 1.0]), np.array([0.125, -0.4, 0.05, 0.7, 0.1]),
          np.array([1.2, 30.0, -10.0, 5.0, 101325.0]), np.array([0.9, 0.0, 0.0, 0.0, 95000.0])]
NORMALS = [np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]), np.array([0.6, 0.0, 0.8]),
           np.array([0.0, -0.28, 0.96])]


@pytest.mark.parametrize("rs", [1, 4, 12])
def test_riemann_consistency(sod_oracle, rs):
    for q in STATES:
        for n in NORMALS:
            F = sod_oracle.riemann(rs, q, q, n)
            Fp = phys_flux(q, n)
            scale = np.abs(Fp).max() + q[4]
            assert np.allclose(F, Fp, rtol=0, atol=1e-13 * scale), (rs, q, n, F, Fp)


@pytest.mark.parametrize("rs", [1, 4, 12])
def test_riemann_supersonic_upwind(sod_oracle, rs):
    # every wave runs to the right -> the flux is the left physical flux (Roe property for rs 12)
    gam = 1004.5 / (1004.5 - 8314.46261815324 / 28.970418)
    qL = np.array([1.0, 3.0 * np.sqrt(gam), 0.1, -0.2, 1.0])
    qR = np.array([0.5, 2.5 * np.sqrt(gam * 0.4 / 0.5), 0.3, 0.0, 0.4])
    n = np.array([1.0, 0.0, 0.0])
    F = sod_oracle.riemann(rs, qL, qR, n)
    Fp = phys_flux(qL, n)
    assert np.allclose(F, Fp, rtol=1e-12, atol=1e-12 * np.abs(Fp).max()), (rs, F, Fp)


@pytest.mark.parametrize("rs", [1, 12])
def test_riemann_stationary_contact(sod_oracle, rs):
    qL = np.array([1.0, 0.0, 0.0, 0.0, 1.0])
    qR = np.array([0.125, 0.0, 0.0, 0.0, 1.0])
    F = sod_oracle.riemann(rs, qL, qR, np.array([1.0, 0.0, 0.0]))
    assert abs(F[0]) < 1e-14 and abs(F[4]) < 1e-14
    assert abs(F[1] - 1.0) < 1e-14


@pytest.mark.parametrize("rs", [1, 4, 12])
def test_riemann_mirror_symmetry(sod_oracle, rs):
    # reflect the problem: swap the states, negate the normal velocity and the normal -> mass flux changes sign
    qL = np.array([1.0, 0.3, 0.2, 0.0, 1.0])
    qR = np.array([0.4, -0.1, 0.2, 0.0, 0.5])
    n = np.array([1.0, 0.0, 0.0])
    F = sod_oracle.riemann(rs, qL, qR, n)
    mL = qR.copy(); mL[1] = -mL[1]
    mR = qL.copy(); mR[1] = -mR[1]
    G = sod_oracle.riemann(rs, mL, mR, n)
    assert np.allclose(G[0], -F[0], atol=1e-14)
    assert np.allclose(G[1], F[1], atol=1e-14)      # normal momentum flux is even
    assert np.allclose(G[4], -F[4], atol=1e-13)


def roe_textbook(qL, qR, n, gam, R):
    """Roe's flux by a numerical eigen-decomposition of the Roe-averaged flux Jacobian (ideal gas)."""
    def cons(q):
        r, u, v, w, p = q
        return np.array([r, r * u, r * v, r * w, p / (gam - 1) + 0.5 * r * (u * u + v * v + w * w)])

    rl, rr = qL[0], qR[0]
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    HL = gam / (gam - 1) * qL[4] / rl + 0.5 * np.dot(qL[1:4], qL[1:4])
    HR = gam / (gam - 1) * qR[4] / rr + 0.5 * np.dot(qR[1:4], qR[1:4])
    vel = (sl * qL[1:4] + sr * qR[1:4]) / (sl + sr)
    H = (sl * HL + sr * HR) / (sl + sr)
    un = vel @ n
    q2 = vel @ vel
    g1 = gam - 1.0
    A = np.zeros((5, 5))   # dF/dU along n
    A[0, 1:4] = n
    for a in range(3):
        A[1 + a, 0] = g1 * 0.5 * q2 * n[a] - vel[a] * un
        A[1 + a, 1:4] = vel[a] * n - g1 * vel * n[a]
        A[1 + a, 1 + a] += un
        A[1 + a, 4] = g1 * n[a]
    A[4, 0] = un * (g1 * 0.5 * q2 - H)
    A[4, 1:4] = H * n - g1 * un * vel
    A[4, 4] = gam * un
    lam, Rv = np.linalg.eig(A)
    absA = (Rv @ np.diag(np.abs(lam)) @ np.linalg.inv(Rv)).real
    FL = phys_flux(qL, n, cp=gam / g1 * R, R=R)
    FR = phys_flux(qR, n, cp=gam / g1 * R, R=R)
    return 0.5 * (FL + FR) - 0.5 * absA @ (cons(qR) - cons(qL))


def test_lmroe_is_textbook_roe(sod_oracle):
    R = 8314.46261815324 / 28.970418
    gam = 1004.5 / (1004.5 - R)
    rng = np.random.default_rng(7)
    for _ in range(50):
        qL = np.array([rng.uniform(0.2, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.2, 2)])
        qR = np.array([rng.uniform(0.2, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.2, 2)])
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        F = sod_oracle.riemann(12, qL, qR, n)
        G = roe_textbook(qL, qR, n, gam, R)
        assert np.allclose(F, G, rtol=1e-10, atol=1e-11), (F, G)


def test_sound_speed_closed_form(sod_oracle):
    R = 8314.46261815324 / 28.970418
    gam = 1004.5 / (1004.5 - R)
    for r, p in [(1.0, 1.0), (0.125, 0.1), (1.177, 101325.0)]:
        a = sod_oracle.sound_speed([r], p)
        assert abs(a - np.sqrt(gam * p / r)) <= 1e-14 * a


def test_muscl_constant_and_linear(sod_oracle):
    n = np.array([1.0, 0.0, 0.0])
    q = np.array([1.1, 0.2, 0.0, 0.0, 1.3])
    dlc = np.full(4, 0.01)
    res = sod_oracle.conv_face(q, q, q, q, dlc, n, 0.5)
    F = sod_oracle.riemann(1, q, q, n) * 0.5
    assert np.array_equal(res[1:], F[1:]) and abs(res[0] - F[0]) <= 1e-16 * abs(F[0]) + 1e-300
    # linear field, equal spacings: the limited slope equals the slope (minmod) and both sides meet at the face
    base = np.array([1.0, 0.1, 0.0, 0.0, 1.0]); grad = np.array([0.01, 0.002, 0.0, 0.0, 0.03])
    cells = [base + grad * (m - 0.5) for m in (-1, 0, 1, 2)]
    res = sod_oracle.conv_face(*cells, dlc, n, 1.0)
    qface = base + grad * 0.5 - grad * 0.5   # value at the face between cells 0 and 1 (x = 0.5 cell units)
    qface = 0.5 * (cells[1] + cells[2])
    F = sod_oracle.riemann(1, qface, qface, n)
    assert np.allclose(res, F, rtol=1e-13, atol=1e-15)
