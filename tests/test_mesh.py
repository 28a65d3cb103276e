"""The lattice mesh builder (include/richmesh.h) against an independent
Delaunay triangulation (scipy/Qhull) and the VoronoiMesh conventions it
follows."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

import cases
from rich_amd import mesh as rmesh


def _delaunay_faces(m):
    """Voronoi faces from Qhull's Delaunay of all points (real + ghosts):
    pairs with at least one real point whose Voronoi edge is longer than
    1e-8 sqrt(R1 R2) (VoronoiMesh.cpp:287-288)."""
    P = np.column_stack([m.pt_x, m.pt_y])
    tri = Delaunay(P)
    S = tri.simplices
    A, B, C = P[S[:, 0]], P[S[:, 1]], P[S[:, 2]]
    b, c = B - A, C - A
    d = 2 * (b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0])
    b2, c2 = (b ** 2).sum(1), (c ** 2).sum(1)
    ux = (c[:, 1] * b2 - b[:, 1] * c2) / d
    uy = (b[:, 0] * c2 - c[:, 0] * b2) / d
    cc = A + np.column_stack([ux, uy])
    R = np.hypot(ux, uy)
    faces = set()
    for t in range(len(S)):
        for k in range(3):
            nb = tri.neighbors[t, k]
            if nb < 0 or nb < t:
                continue
            p, q = S[t, (k + 1) % 3], S[t, (k + 2) % 3]
            if p >= m.n_cells and q >= m.n_cells:
                continue
            if np.hypot(*(cc[t] - cc[nb])) > 1e-8 * np.sqrt(R[t] * R[nb]):
                faces.add((min(p, q), max(p, q)))
    return faces


@pytest.mark.parametrize("bc", [rmesh.RIGID, rmesh.PERIODIC])
@pytest.mark.parametrize("jitter", [1e-3, 3e-2])
def test_faces_match_delaunay(bc, jitter):
    m = rmesh.lattice(17, 11, (0.0, 1.7, -0.5, 0.6), jitter=jitter, seed=13, bc=bc)
    mine = {(min(a, b), max(a, b)) for a, b in zip(m.face_first, m.face_second)}
    ref = _delaunay_faces(m)
    # faces of a real cell must agree; qhull may add ghost-side faces that
    # touch no real cell's polygon (they are not in the reference either)
    assert mine <= ref
    real_ref = {f for f in ref if f[0] < m.n_cells and f[1] < m.n_cells}
    real_mine = {f for f in mine if f[0] < m.n_cells and f[1] < m.n_cells}
    assert real_mine == real_ref


@pytest.mark.parametrize("bc", [rmesh.RIGID, rmesh.PERIODIC])
def test_geometry_and_csr(bc):
    box = (0.0, 2.0, 0.0, 1.0)
    m = rmesh.lattice(24, 12, box, jitter=1e-3, seed=1, bc=bc)
    N = m.n_cells
    area = (box[1] - box[0]) * (box[3] - box[2])
    assert m.volume.sum() == pytest.approx(area, rel=1e-12)
    assert m.tess_volume.sum() == pytest.approx(area, rel=1e-12)
    np.testing.assert_allclose(m.volume, m.tess_volume, rtol=1e-10)
    # CSR: ascending faces, each face listed by its real sides
    cnt = np.zeros(N, int)
    for i in range(N):
        fs = m.csr_face[m.csr_off[i]:m.csr_off[i + 1]]
        assert np.all(np.diff(fs) > 0)
        assert np.all((m.face_first[fs] == i) | (m.face_second[fs] == i))
        cnt[i] = len(fs)
    real_sides = (m.face_first < N).sum() + (m.face_second < N).sum()
    assert cnt.sum() == real_sides
    # centres of mass inside the cell's bounding box of vertices
    for i in range(0, N, 7):
        fs = m.csr_face[m.csr_off[i]:m.csr_off[i + 1]]
        xs = np.r_[m.v1x[fs], m.v2x[fs]]
        ys = np.r_[m.v1y[fs], m.v2y[fs]]
        assert xs.min() <= m.cm_x[i] <= xs.max() and ys.min() <= m.cm_y[i] <= ys.max()
    # ghost centres of mass follow their originals
    g = np.arange(m.n_ghosts)
    o = m.ghost_orig
    if bc == rmesh.PERIODIC:
        np.testing.assert_allclose(m.cm_x[N + g] - m.pt_x[N + g], m.cm_x[o] - m.pt_x[o], atol=1e-12)
        np.testing.assert_allclose(m.cm_y[N + g] - m.pt_y[N + g], m.cm_y[o] - m.pt_y[o], atol=1e-12)
    else:
        # mirror images: the wall bisects point/ghost and CM/ghost CM alike
        mx = 0.5 * (m.pt_x[N + g] + m.pt_x[o]) - 0.5 * (m.cm_x[N + g] + m.cm_x[o])
        my = 0.5 * (m.pt_y[N + g] + m.pt_y[o]) - 0.5 * (m.cm_y[N + g] + m.cm_y[o])
        assert np.all((np.abs(mx) < 1e-12) | (np.abs(my) < 1e-12))


def test_rigid_boundary_faces_are_mirror_faces_on_walls():
    m = rmesh.lattice(13, 9, (-1.0, 1.0, -0.5, 0.5), jitter=1e-2, seed=3, bc=rmesh.RIGID)
    N = m.n_cells
    bnd = (m.face_first >= N) | (m.face_second >= N)
    assert bnd.any()
    orig_a = m.orig(m.face_first[bnd])
    orig_b = m.orig(m.face_second[bnd])
    assert np.all(orig_a == orig_b)
    on_wall = (np.isclose(m.v1x[bnd], -1) & np.isclose(m.v2x[bnd], -1)) | \
              (np.isclose(m.v1x[bnd], 1) & np.isclose(m.v2x[bnd], 1)) | \
              (np.isclose(m.v1y[bnd], -0.5) & np.isclose(m.v2y[bnd], -0.5)) | \
              (np.isclose(m.v1y[bnd], 0.5) & np.isclose(m.v2y[bnd], 0.5))
    assert np.all(on_wall)


def test_flip_changes_orientation_only():
    a = rmesh.lattice(10, 10, jitter=1e-2, seed=5, bc=rmesh.PERIODIC)
    b = rmesh.lattice(10, 10, jitter=1e-2, seed=5, bc=rmesh.PERIODIC, flip_seed=3)
    assert a.n_faces == b.n_faces
    flipped = a.face_first != b.face_first
    assert 0 < flipped.sum() < a.n_faces
    np.testing.assert_array_equal(a.face_first[flipped], b.face_second[flipped])
    np.testing.assert_array_equal(a.v1x, b.v1x)
    np.testing.assert_array_equal(a.cm_x, b.cm_x)


def test_kh_size():
    m = rmesh.lattice(256, 128, (0.0, 2.0, 0.0, 1.0), jitter=1e-3, seed=1, bc=rmesh.PERIODIC)
    assert m.n_cells == 32768
    # interior faces ~3N (two lattice edges + one diagonal per cell) plus the
    # periodic boundary layer
    assert 3.0 < m.n_faces / m.n_cells < 3.05


@pytest.mark.parametrize("n", [20, 64])
def test_rigid_ghost_touches_only_its_original(n):
    """RigidWallGenerator copies the real cell of a ghost's first boundary
    face (RigidWallGenerator.cpp:5-37).  On these lattices every rigid ghost
    borders exactly one face, shared with the cell it mirrors
    (GetOriginalIndex), so that rule and the mirror rule coincide."""
    m = cases.sedov_mesh(n, flip_seed=3)
    N, G = m.n_cells, m.n_ghosts
    count = np.zeros(G, int)
    for f in range(m.n_faces):
        a, b = m.face_first[f], m.face_second[f]
        for g, r in ((a, b), (b, a)):
            if g >= N:
                count[g - N] += 1
                assert r == m.ghost_orig[g - N]
    assert np.all(count[count > 0] == 1)
