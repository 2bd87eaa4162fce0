"""The oracle's implicit residual smoothing (IRS_apply, src/lib/numerics/time/explicit/Lib_IRS.f90:53-96), pinned by
closed forms.

Per computational direction IRS takes njacobi Jacobi sweeps R* <- (R + beta (R*(c-1) + R*(c+1))) / (1 + 2 beta), from
R* = R, with a ghost layer of R behind every block face: zero gradient (the boundary cell's own value), or, behind a
connection, the donor cell's value.  A discrete Fourier mode is an eigenvector of one sweep: sin(theta i + phi) on a
periodic direction (theta = 2 pi k / n), and cos(theta (i - 1/2)) between zero-gradient faces (theta = pi k / n, the
ghost then equals the boundary cell's value).  The factor after J sweeps follows a_0 = 1,
a_(n+1) = (1 + 2 beta cos(theta) a_n) / (1 + 2 beta), and the directions multiply for a separable mode.  A wrong
ghost (periodic or zero-gradient), a wrong sweep count, beta in the wrong place or an unsmoothed direction each change
the factors.  The ghost layer is filled once per direction, before the sweeps (irs_smooth_blk keeps the ghosts of R*
at R's values), so from the second sweep on the cells next to a block face see the unsmoothed ghost: after J sweeps
the closed form holds everywhere for J = 1 and, for J > 1, on the cells at least J - 1 cells from every block face.
The tests run the primitive update's smoothing (the raw residual); the conservative update's scaling is checked on
its own."""
import numpy as np
import pytest

from mosegpu import cases
from mosegpu.core import BC_EXTRAPOLATION, BC_SYMMETRY, BC_WALL_ADIABATIC
from oracle_bind import OracleCase


def amp(theta, beta, J):
    a = 1.0
    for _ in range(J):
        a = (1.0 + 2.0 * beta * np.cos(theta) * a) / (1.0 + 2.0 * beta)
    return a


def _set_irs(case, beta, J, integration):
    case.scheme.irs = 1
    case.scheme.irs_beta = beta
    case.scheme.irs_njacobi = J
    case.scheme.integration = {"cons": 0, "prim": 1}[integration]
    return case


def _global_mode(case, dims_grid, shape_fn):
    """R (var-last arena) with every flow variable s equal to (s + 1) * shape(I, J, K) over the lattice's global
    1-based cell indices, ghosts filled with 1e3 (IRS must overwrite the ones it reads)."""
    R = np.full_like(case.P, 1e3)
    nbx, nby, nbz, ox, oy, oz = dims_grid
    for b in range(case.layout.nb):
        a_, r = b % nbx, b // nbx
        b_, c_ = r % nby, r // nby
        V = case.layout.interior(R, b)
        n1, n2, n3 = case.layout.dims[b]
        K, J, I = np.meshgrid(np.arange(1, n3 + 1) + oz[c_], np.arange(1, n2 + 1) + oy[b_],
                              np.arange(1, n1 + 1) + ox[a_], indexing="ij")
        for s in range(V.shape[0]):
            V[s] = (s + 1) * shape_fn(I, J, K)
    return R


def _one_block(case):
    return (1, 1, 1, [0], [0], [0])


def _far(dims, J):
    """mask of a block's cells at least J - 1 cells from each of its faces (indices k, j, i of the interior view)"""
    n1, n2, n3 = dims
    m = lambda n: (np.arange(1, n + 1) >= J) & (np.arange(1, n + 1) <= n - J + 1)
    return m(n3)[:, None, None] & m(n2)[None, :, None] & m(n1)[None, None, :]


def _check(c, R, before, g, J, nflow, tol=1e-14):
    n_far = 0
    for b in range(c.layout.nb):
        A, B0 = c.layout.interior(R, b)[:nflow], before[b][:nflow]
        mask = _far(c.layout.dims[b], J)
        n_far += int(mask.sum())
        d = np.abs(A - g * B0)[:, mask]
        assert d.size == 0 or np.max(d) <= tol * np.max(np.abs(B0)), (b, np.max(d))
    return n_far


@pytest.mark.parametrize("beta,J", [(0.5, 1), (0.8, 1), (0.5, 2), (0.8, 3)])
def test_irs_periodic_fourier_mode(beta, J):
    n = (12, 10, 8)
    c = _set_irs(cases.box_case(n, {f: "periodic" for f in range(1, 7)}, viscous=False), beta, J, "prim")
    th = [2 * np.pi * 1 / n[0], 2 * np.pi * 2 / n[1], 2 * np.pi * 1 / n[2]]
    ph = (0.3, 0.1, 0.7)
    f = lambda I, J_, K: np.sin(th[0] * I + ph[0]) * np.sin(th[1] * J_ + ph[1]) * np.sin(th[2] * K + ph[2])
    R = _global_mode(c, _one_block(c), f)
    before = [c.layout.interior(R, 0).copy()]
    OracleCase(c, nthreads=1).irs(R, 1.0)
    g = amp(th[0], beta, J) * amp(th[1], beta, J) * amp(th[2], beta, J)
    assert 0.05 < g < 0.95
    assert _check(c, R, before, g, J, 5) > 0
    if J > 1:     # the stale ghosts are the reference's: next to a face the factor is not the closed form
        A = c.layout.interior(R, 0)
        assert np.max(np.abs(A - g * before[0])) > 1e-6 * np.max(np.abs(before[0]))


@pytest.mark.parametrize("J", [1, 2])
def test_irs_zero_gradient_faces_cosine_mode(J):
    n = (9, 7, 6)
    faces = {1: BC_EXTRAPOLATION, 2: BC_WALL_ADIABATIC, 3: BC_SYMMETRY, 4: BC_EXTRAPOLATION, 5: BC_SYMMETRY,
             6: BC_WALL_ADIABATIC}
    c = _set_irs(cases.box_case(n, faces, viscous=True), 0.6, J, "prim")
    th = [np.pi * 2 / n[0], np.pi * 1 / n[1], np.pi * 3 / n[2]]
    f = lambda I, J_, K: np.cos(th[0] * (I - 0.5)) * np.cos(th[1] * (J_ - 0.5)) * np.cos(th[2] * (K - 0.5))
    R = _global_mode(c, _one_block(c), f)
    before = [c.layout.interior(R, 0).copy()]
    OracleCase(c, nthreads=1).irs(R, 1.0)
    g = amp(th[0], 0.6, J) * amp(th[1], 0.6, J) * amp(th[2], 0.6, J)
    assert _check(c, R, before, g, J, 5) > 0


@pytest.mark.parametrize("J", [1, 2])
def test_irs_on_a_block_lattice_is_the_global_mode(J):
    """2 x 2 x 2 blocks of uneven widths: connections inside, a wall and far field outside.  Over the lattice's global
    cell indices the cosine mode is an eigenvector of one sweep (a connection's ghost is its donor's value); the
    Spalart-Allmaras variable is not smoothed."""
    c = _set_irs(cases.channel_lattice(nb=(2, 2, 2), wmin=4, wmax=9, seed=3), 0.5, J, "prim")
    dims = c.layout.dims
    wx = [dims[a][0] for a in range(2)]
    wy = [dims[2 * b][1] for b in range(2)]
    wz = [dims[4 * k][2] for k in range(2)]
    N = (sum(wx), sum(wy), sum(wz))
    grid = (2, 2, 2, [0, wx[0]], [0, wy[0]], [0, wz[0]])
    th = [np.pi * 1 / N[0], np.pi * 2 / N[1], np.pi * 1 / N[2]]
    f = lambda I, J_, K: np.cos(th[0] * (I - 0.5)) * np.cos(th[1] * (J_ - 0.5)) * np.cos(th[2] * (K - 0.5))
    R = _global_mode(c, grid, f)
    nt = c.scheme.nsc + 4
    before = [c.layout.interior(R, b).copy() for b in range(c.layout.nb)]
    OracleCase(c, nthreads=1).irs(R, 1.0)
    g = amp(th[0], 0.5, J) * amp(th[1], 0.5, J) * amp(th[2], 0.5, J)
    assert _check(c, R, before, g, J, nt) > 0
    for b in range(c.layout.nb):
        assert np.array_equal(c.layout.interior(R, b)[nt], before[b][nt])


def test_irs_conservative_scaling():
    """With the conservative update the residual is first scaled to -R / vol * dt * strang (every variable); a
    uniform field on a uniform periodic box then stays uniform through the sweeps."""
    n = (6, 5, 4)
    c = _set_irs(cases.box_case(n, {f: "periodic" for f in range(1, 7)}, viscous=False), 0.5, 2, "cons")
    R = _global_mode(c, _one_block(c), lambda I, J_, K: np.ones_like(I, dtype=np.float64))
    vol = c.layout.interior(c.geom.vol[None, :], 0)[0]
    dt = 2.5e-4
    OracleCase(c, nthreads=1).irs(R, dt)
    A = c.layout.interior(R, 0)
    for s in range(A.shape[0]):
        want = -(s + 1) / vol * dt
        assert np.max(np.abs(A[s] - want)) <= 4e-16 * np.max(np.abs(want)), s
