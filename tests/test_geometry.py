"""Pins of the host geometry builder (mose_metrics_build: Compute_Norm_Area and Compute_Metric_Tensor,
Lib_Metrics.f90:199-510, and the ghost metrics of :60-102) against exact values, on meshes whose cells are
parallelepipeds: an affine block x = x0 + J (i, j, k) with a full, non-symmetric J, and a block stretched along x and
sheared in y and z.

For a parallelepiped with edge vectors e_i, e_j, e_k (columns of J), the reference's formulas are exact:
- A(h, :) = e_h (the mean of the four edges along h), so dl_h = |e_h|, the metric M = A^-T = J^-1
  (M(a, b) = d xi_a / d x_b, the stored m[(a + 3 b) NC + row]) and the five-tetrahedra volume = det J;
- the face area vector (half the cross product of the face diagonals) is e_j x e_k on i-faces, e_k x e_i on
  j-faces and e_i x e_j on k-faces, pointing towards the larger index.
Ghost cells: on an affine block the cubic node extrapolation (walls, far field, extrapolation) and the translated
donor nodes of a periodic connection both reproduce the affine map, so every ghost carries the interior values.
A transposed metric, a dropped term of the cofactors or a sign of a normal fails these by O(1).
"""
import numpy as np
import pytest

from mosegpu import cases
from mosegpu.core import BC_CONNECTION, BC_EXTRAPOLATION, BC_WALL_ADIABATIC, GC, Layout, metrics_build

J_AFFINE = np.array([[0.31, 0.05, -0.07],
                     [0.04, 0.22, 0.03],
                     [-0.02, 0.06, 0.47]])


def affine_nodes(dims, J, x0=(0.3, -0.2, 0.1)):
    n1, n2, n3 = dims
    k, j, i = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    idx = np.stack([i, j, k], axis=-1).astype(np.float64)
    return np.asarray(x0) + idx @ J.T          # (n3+1, n2+1, n1+1, 3)


def build(dims, nodes, faces, ndir=3):
    lay = Layout([dims])
    rec = cases.block_records(0, dims, faces, {f: v for f, v in _periodic_donors(dims).items()
                                               if faces[f] == BC_CONNECTION})
    g = metrics_build(lay, [np.ascontiguousarray(nodes)], np.ascontiguousarray(rec), ndir)
    return lay, g


def _periodic_donors(dims):
    n1, n2, n3 = dims
    return {1: (0, 2, lambda i, j, k: (np.full_like(i, n1), j, k), True),
            2: (0, 1, lambda i, j, k: (np.full_like(i, 1), j, k), True),
            3: (0, 4, lambda i, j, k: (i, np.full_like(j, n2), k), True),
            4: (0, 3, lambda i, j, k: (i, np.full_like(j, 1), k), True),
            5: (0, 6, lambda i, j, k: (i, j, np.full_like(k, n3)), True),
            6: (0, 5, lambda i, j, k: (i, j, np.full_like(k, 1)), True)}


def cell_views(lay, g):
    dl = lay.block_view(g.dl.reshape(3, -1), 0)
    m = lay.block_view(g.m.reshape(9, -1), 0)
    vol = lay.block_view(g.vol.reshape(1, -1), 0)[0]
    return dl, m, vol


def face_views(dims, g):
    """fm1..3 as [4, k, j, i] arrays over each direction's faces (n, then area)."""
    n1, n2, n3 = dims
    f1 = g.fm1.reshape(4, n3, n2, n1 + 1)
    f2 = g.fm2.reshape(4, n3, n2 + 1, n1)
    f3 = g.fm3.reshape(4, n3 + 1, n2, n1)
    return f1, f2, f3


def check_cell(dl_c, m_c, vol_c, E, rtol=1e-13):
    """dl, m (9,) and vol of one cell against its edge vectors E (columns e_i, e_j, e_k)."""
    assert np.allclose(dl_c, np.linalg.norm(E, axis=0), rtol=rtol, atol=0)
    Minv = np.linalg.inv(E)                                   # M(a, b) = d xi_a / d x_b
    got = np.array(m_c).reshape(3, 3, order="F")               # m[(a + 3 b)] = M(a, b)
    assert np.allclose(got, Minv, rtol=0, atol=rtol * np.abs(Minv).max())
    assert np.isclose(vol_c, np.linalg.det(E), rtol=rtol, atol=0)


def check_face(fm, avec, rtol=1e-13):
    area = np.linalg.norm(avec)
    assert np.isclose(fm[3], area, rtol=rtol, atol=0)
    assert np.allclose(fm[:3], avec / area, rtol=0, atol=rtol)


@pytest.mark.parametrize("faces", ["periodic", "extrapolation", "walls"])
def test_affine_block_exact_geometry(faces):
    dims = (5, 4, 3)
    E = J_AFFINE
    if faces == "periodic":
        ft = {f: BC_CONNECTION for f in range(1, 7)}
    elif faces == "extrapolation":
        ft = {f: BC_EXTRAPOLATION for f in range(1, 7)}
    else:
        ft = {f: BC_WALL_ADIABATIC for f in range(1, 7)}
    lay, g = build(dims, affine_nodes(dims, E), ft)
    dl, m, vol = cell_views(lay, g)
    n1, n2, n3 = dims
    # every interior cell, and every ghost of the two layers behind each face (the ghost rows the kernels read)
    cells = [(i, j, k) for k in range(n3) for j in range(n2) for i in range(n1)]
    for f in range(6):
        ax, hi = f // 2, f % 2
        for g_ in range(1, GC + 1):
            for k in range(n3):
                for j in range(n2):
                    for i in range(n1):
                        c = [i, j, k]
                        if c[ax] != (dims[ax] - 1 if hi else 0):
                            continue
                        c[ax] = c[ax] + g_ if hi else c[ax] - g_
                        cells.append(tuple(c))
    for (i, j, k) in cells:
        r = (k + GC, j + GC, i + GC)
        check_cell(dl[(slice(None),) + r], m[(slice(None),) + r], vol[r], E)
    f1, f2, f3 = face_views(dims, g)
    e = [E[:, 0], E[:, 1], E[:, 2]]
    a1, a2, a3 = np.cross(e[1], e[2]), np.cross(e[2], e[0]), np.cross(e[0], e[1])
    for fm, av in ((f1, a1), (f2, a2), (f3, a3)):
        flat = fm.reshape(4, -1)
        for q in range(flat.shape[1]):
            check_face(flat[:, q], av)


def test_stretched_sheared_block_exact_geometry():
    """x = f(i) + a j + b k, y = c j, z = d k with a geometric f: every cell is a parallelepiped with edges
    (f(i) - f(i-1), 0, 0), (a, c, 0), (b, 0, d)."""
    dims = (6, 3, 4)
    n1, n2, n3 = dims
    a, b, c_, d = 0.07, -0.05, 0.3, 0.2
    fi = np.concatenate([[0.0], np.cumsum(0.1 * 1.4 ** np.arange(n1))])
    k, j, i = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    nodes = np.stack([fi[i] + a * j + b * k, c_ * j, d * k], axis=-1).astype(np.float64)
    lay, g = build(dims, nodes, {f: BC_EXTRAPOLATION for f in range(1, 7)})
    dl, m, vol = cell_views(lay, g)
    f1, f2, f3 = face_views(dims, g)
    ej, ek = np.array([a, c_, 0.0]), np.array([b, 0.0, d])
    for ii in range(n1):
        ei = np.array([fi[ii + 1] - fi[ii], 0.0, 0.0])
        E = np.stack([ei, ej, ek], axis=1)
        for kk in range(n3):
            for jj in range(n2):
                r = (kk + GC, jj + GC, ii + GC)
                check_cell(dl[(slice(None),) + r], m[(slice(None),) + r], vol[r], E)
                check_face(f2[:, kk, jj, ii], np.cross(ek, ei))
                check_face(f2[:, kk, jj + 1, ii], np.cross(ek, ei))
                check_face(f3[:, kk, jj, ii], np.cross(ei, ej))
        for kk in range(n3):
            for jj in range(n2):
                check_face(f1[:, kk, jj, ii], np.cross(ej, ek))


def test_degenerate_cell_is_refused():
    """A cell of zero volume (det = 0) is MOSE_ERR_MESH, as Compute_Metric_Tensor stops on det = 0."""
    dims = (2, 2, 2)
    nodes = affine_nodes(dims, J_AFFINE)
    nodes[:, :, 1, :] = nodes[:, :, 0, :]          # collapse the first column of cells
    with pytest.raises(Exception):
        build(dims, nodes, {f: BC_EXTRAPOLATION for f in range(1, 7)})


def brute_wall_distance(lay, nodes, rec):
    """Compute_Yn by a scan of every wall-face centroid (the k-d tree's answer must be this one)."""
    from mosegpu.core import BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL
    cen = []
    for r in rec:
        if r["typ"] not in (BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL):
            continue
        nd = nodes[r["b"]]
        i, j, k, f = int(r["i"]), int(r["j"]), int(r["k"]), int(r["f"])
        m2 = f % 2
        if f <= 2:
            c = [(i - m2, j - 1, k - 1), (i - m2, j, k - 1), (i - m2, j - 1, k), (i - m2, j, k)]
        elif f <= 4:
            c = [(i - 1, j - m2, k - 1), (i, j - m2, k - 1), (i - 1, j - m2, k), (i, j - m2, k)]
        else:
            c = [(i - 1, j - 1, k - m2), (i, j - 1, k - m2), (i - 1, j, k - m2), (i, j, k - m2)]
        p = [nd[kk, jj, ii] for ii, jj, kk in c]
        cen.append([0.25 * (p[0][a] + p[1][a] + p[2][a] + p[3][a]) for a in range(3)])
    cen = np.array(cen)
    out = []
    for b, nd in enumerate(nodes):
        order = [(0, 0, 0), (0, -1, 0), (0, -1, -1), (0, 0, -1), (-1, 0, 0), (-1, -1, 0), (-1, -1, -1), (-1, 0, -1)]
        n3, n2, n1 = nd.shape[0] - 1, nd.shape[1] - 1, nd.shape[2] - 1
        s = None
        for (di, dj, dk) in order:
            v = nd[1 + dk:n3 + 1 + dk, 1 + dj:n2 + 1 + dj, 1 + di:n1 + 1 + di]
            s = v.copy() if s is None else s + v
        ctr = 1.0 / 8.0 * s                                    # (n3, n2, n1, 3)
        dx = cen[None, None, None, :, 0] - ctr[..., 0:1]
        dy = cen[None, None, None, :, 1] - ctr[..., 1:2]
        dz = cen[None, None, None, :, 2] - ctr[..., 2:3]
        out.append(np.sqrt((dx * dx + dy * dy + dz * dz).min(axis=-1)))
    return out


def test_wall_distance_is_the_nearest_wall_centroid():
    """mose_wall_distance on the channel lattice of seeded uneven blocks (wall at y = 0), its nodes sheared: the
    k-d tree gives the scan's distances bit for bit; on the unsheared lattice the distance is the cell centre's y."""
    from mosegpu.core import wall_distance
    case = cases.channel_lattice(nb=(2, 2, 2), wmin=3, wmax=6, seed=11)
    lay = case.layout
    yn = np.zeros(lay.NC)
    wall_distance(lay, case.nodes, case.records, yn)
    for b, nd in enumerate(case.nodes):
        ctr = 0.125 * (nd[:-1, :-1, :-1] + nd[1:, :-1, :-1] + nd[:-1, 1:, :-1] + nd[:-1, :-1, 1:] + nd[1:, 1:, :-1]
                       + nd[1:, :-1, 1:] + nd[:-1, 1:, 1:] + nd[1:, 1:, 1:])
        assert np.allclose(lay.interior(yn, b)[0] if lay.interior(yn, b).ndim == 4 else lay.interior(yn, b),
                           ctr[..., 1], rtol=1e-14, atol=0)
    sheared = [nd + np.stack([0.3 * nd[..., 1] + 0.1 * nd[..., 2], np.zeros_like(nd[..., 1]),
                              0.2 * nd[..., 0] * nd[..., 1]], axis=-1) for nd in case.nodes]
    yn2 = np.zeros(lay.NC)
    wall_distance(lay, sheared, case.records, yn2)
    ref = brute_wall_distance(lay, sheared, case.records)
    for b in range(lay.nb):
        got = lay.interior(yn2.reshape(1, -1), b)[0]
        assert np.array_equal(got, ref[b]), np.max(np.abs(got - ref[b]))


def test_wall_distance_connection_ghosts():
    """Yn_Connection: the first ghost layer behind a connection face holds the donor cell's distance."""
    from mosegpu.core import BC_CONNECTION, wall_distance
    case = cases.channel_lattice(nb=(2, 2, 2), wmin=3, wmax=6, seed=11)
    lay = case.layout
    yn = np.full(lay.NC, -1.0)
    wall_distance(lay, case.nodes, case.records, yn)
    guide = {1: (1, 0, 0), 2: (-1, 0, 0), 3: (0, 1, 0), 4: (0, -1, 0), 5: (0, 0, 1), 6: (0, 0, -1)}
    nconn = 0
    for r in case.records:
        if r["typ"] != BC_CONNECTION:
            continue
        g = guide[int(r["f"])]
        vb = lay.block_view(yn.reshape(1, -1), int(r["b"]))[0]
        vs = lay.block_view(yn.reshape(1, -1), int(r["bs"]))[0]
        ghost = vb[r["k"] - g[2] + GC - 1, r["j"] - g[1] + GC - 1, r["i"] - g[0] + GC - 1]
        donor = vs[r["ks"] + GC - 1, r["js"] + GC - 1, r["is_"] + GC - 1]
        assert ghost == donor
        nconn += 1
    assert nconn > 0
