"""Pins of the oracle's symmetry faces (MOSE 300, internal type 3: Ghost_Symmetry, BC_Symmetry_Eul and
BC_Symmetry_Visc) on the CPU.

Take a periodic box [0, 2L] along one axis and a state that is mirror-symmetric about x = L: density, pressure, nu~
and the transverse velocities even, the normal velocity odd (the second half is the first half reversed, the normal
velocity negated).  The box then stays symmetric.  Its first half must be what a box [0, L] gives with symmetry
faces at x = 0 and x = L.  There:
- the mirror ghost is the full box's neighbour cell;
- the pressure flux of BC_Symmetry_Eul is the full box's face flux, whose mass and energy parts vanish;
- BC_Symmetry_Visc's gradients (the cell and its mirror image, and the four transverse neighbours with theirs) are
  the full box's face gradients.
The two runs differ only in rounding.  A velocity mirrored with the wrong sign, a transverse gradient term dropped
or a stencil one cell off gives differences many orders above it.

The second ghost layer behind a symmetry face is never written (Ghost_Symmetry fills the first only).  The half box
starts with NaN there: the run must stay finite.
"""
import numpy as np
import pytest

from mosegpu import cases
from mosegpu.core import BC_NONE, BC_SYMMETRY, BC_WALL_ADIABATIC, GC
from oracle_bind import OracleCase


def _faces(axis, sym, walls):
    f = {q: "periodic" for q in range(1, 7)}
    f[2 * axis + 1] = f[2 * axis + 2] = BC_SYMMETRY if sym else "periodic"
    if walls is not None:
        f[2 * walls + 1] = f[2 * walls + 2] = BC_WALL_ADIABATIC
    return f


def stretched_nodes(dims, ext, axis, half_n):
    """Nodes of a box whose spacing along `axis` grows geometrically (ratio 1.35) from x = 0 over the first half_n
    cells and mirrors about x = L after them (a full box of 2 half_n cells), uniform along the other axes."""
    n1, n2, n3 = dims
    w = 1.35 ** np.arange(half_n)
    xh = np.concatenate([[0.0], np.cumsum(w)]) / w.sum() * ext[axis]
    x = xh if dims[axis] == half_n else np.concatenate([xh, 2 * ext[axis] - xh[-2::-1]])
    axes = [np.arange(n + 1, dtype=np.float64) / n * e for n, e in zip((n1, n2, n3), ext)]
    axes[axis] = x
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([X, Y, Z], axis=-1)


def mirror_pair(axis, half_dims, viscous, nrans=0, walls=None, riemann="HLLC", limiter="vanleer", seed=5,
                stretch=False):
    """The half box with symmetry faces on `axis` and the full periodic box, both at the same mirror-symmetric state.
    stretch: the cells grow along `axis` away from x = 0 and the full box is the half box and its mirror image."""
    full_dims = list(half_dims)
    full_dims[axis] *= 2
    ext = [1.0, 0.8, 0.6]
    full_ext = list(ext)
    full_ext[axis] *= 2
    kw = dict(nrans=nrans, viscous=viscous, riemann=riemann, limiter=limiter, ndir=3,
              mu_of_T=cases.sutherland(1.8e-3, 300.0) if viscous else None)
    nh = nf = None
    if stretch:
        nh = stretched_nodes(half_dims, ext, axis, half_dims[axis])
        nf = stretched_nodes(full_dims, ext, axis, half_dims[axis])
    half = cases.box_case(tuple(half_dims), _faces(axis, True, walls), extent=tuple(ext), nodes=nh, **kw)
    full = cases.box_case(tuple(full_dims), _faces(axis, False, walls), extent=tuple(full_ext), nodes=nf, **kw)
    rng = np.random.Generator(np.random.PCG64(seed))
    n3, n2, n1 = half_dims[2], half_dims[1], half_dims[0]
    shape = (n3, n2, n1)
    V = half.layout.interior(half.P, 0)
    V[0] = 1.0 + 0.2 * rng.uniform(-1, 1, shape)
    for q in range(3):
        V[1 + q] = 40.0 * rng.uniform(-1, 1, shape)
    V[4] = 1.0e5 * (1.0 + 0.1 * rng.uniform(-1, 1, shape))
    if nrans:
        V[5] = 1e-4 * (1.0 + rng.uniform(0, 1, shape))
    cases.fill_ghosts_nearest(half.layout, half.P)
    ax = 3 - axis                                   # array axis of the mirrored direction in [var, k, j, i]
    W = full.layout.interior(full.P, 0)
    lo = [slice(None)] * 4
    hi = [slice(None)] * 4
    lo[ax] = slice(0, half_dims[axis])
    hi[ax] = slice(half_dims[axis], None)
    W[tuple(lo)] = V
    mir = np.flip(V, axis=ax).copy()
    mir[1 + axis] = -mir[1 + axis]
    W[tuple(hi)] = mir
    cases.fill_ghosts_nearest(full.layout, full.P)
    # the second ghost layer behind both symmetry faces: never read
    B = half.layout.block_view(half.P, 0)
    for g in (0, B.shape[ax] - 1):
        idx = [slice(None)] * 4
        idx[ax] = g
        B[tuple(idx)] = np.nan
    return half, full


def _compare(half, full, axis, nsteps):
    Ph, Pf = half.P.copy(), full.P.copy()
    OracleCase(half).step(Ph, nsteps)
    OracleCase(full).step(Pf, nsteps)
    V = half.layout.interior(Ph, 0)
    W = full.layout.interior(Pf, 0)
    ax = 3 - axis
    n = V.shape[ax]
    lo = [slice(None)] * 4
    lo[ax] = slice(0, n)
    hi = [slice(None)] * 4
    hi[ax] = slice(n, None)
    A = W[tuple(lo)]
    Bm = np.flip(W[tuple(hi)], axis=ax).copy()
    Bm[1 + axis] = -Bm[1 + axis]
    assert np.all(np.isfinite(V))
    out = []
    for s in range(V.shape[0]):
        scale = np.max(np.abs(A[s]))
        out.append(max(np.max(np.abs(V[s] - A[s])), np.max(np.abs(V[s] - Bm[s]))) / scale)
    return np.array(out)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("viscous", [False, True])
def test_symmetry_face_is_a_mirror(axis, viscous):
    dims = [5, 4, 3]
    half, full = mirror_pair(axis, dims, viscous, riemann="HLLC" if viscous else "LMRoe")
    err = _compare(half, full, axis, 3)
    assert np.max(err) <= 1e-12, err


def test_symmetry_face_is_a_mirror_sa():
    """With Spalart-Allmaras: the eddy viscosity of BC_Symmetry_Visc and the SA source next to the face (walls on
    the y faces give the wall distance)."""
    half, full = mirror_pair(0, [5, 4, 3], True, nrans=1, walls=1)
    err = _compare(half, full, 0, 3)
    assert np.max(err) <= 1e-12, err


def test_symmetry_mistakes_are_seen():
    """The same comparison with the symmetry faces turned into no-flux faces (type 0) fails by far more than the
    bar: the pin can tell a wrong boundary from rounding."""
    half, full = mirror_pair(0, [5, 4, 3], True)
    half.records["typ"] = np.where(half.records["typ"] == BC_SYMMETRY, BC_NONE, half.records["typ"])
    cases.fill_ghosts_nearest(half.layout, half.P)
    err = _compare(half, full, 0, 3)
    assert np.max(err) > 1e-6, err


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("viscous", [False, True])
def test_symmetry_face_is_a_mirror_on_a_stretched_mesh(axis, viscous):
    """The mirror check on cells that grow by 1.35 from face to face along the mirrored axis: the ghosts behind a
    symmetry face carry the metric, lengths and volume of their mirror images (BC_Symmetry_Metrics,
    Lib_Metrics.f90:1044-1102), as the full box's cells across the mirror plane do.  Extrapolated ghost lengths would
    change the limited slopes of the first faces."""
    half, full = mirror_pair(axis, [6, 5, 4], viscous, riemann="HLLC" if viscous else "LMRoe", stretch=True)
    err = _compare(half, full, axis, 3)
    assert np.max(err) <= 1e-12, err


def test_symmetry_ghost_metrics_are_mirror_images():
    """Ghost g behind each symmetry face holds interior cell g - 1's lengths, metric and volume (g = 1, 2), not the
    cubic node extrapolation of the other boundary types."""
    half, _ = mirror_pair(0, [6, 5, 4], False, stretch=True)
    lay, g = half.layout, half.geom
    dl = lay.block_view(g.dl.reshape(3, -1), 0)            # [3, k, j, i] padded
    m = lay.block_view(g.m.reshape(9, -1), 0)
    vol = lay.block_view(g.vol.reshape(1, -1), 0)
    n1 = half.layout.dims[0][0]
    for q in (1, 2):                            # ghost q before the first cell and after the last
        lo, src_lo = GC - q, GC + q - 1
        hi, src_hi = GC + n1 - 1 + q, GC + n1 - q
        for arr in (dl, m, vol):
            assert np.array_equal(arr[:, GC:-GC, GC:-GC, lo], arr[:, GC:-GC, GC:-GC, src_lo])
            assert np.array_equal(arr[:, GC:-GC, GC:-GC, hi], arr[:, GC:-GC, GC:-GC, src_hi])
    # the stretched spacing is not what an extrapolation gives: the first ghost is as long as the first cell
    assert dl[0, GC, GC, GC - 1] != dl[0, GC, GC, GC + 1]
