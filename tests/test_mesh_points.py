"""The point-set mesh builder (rich_amd.mesh.points: Qhull Delaunay of the
points and their wall reflections, then rm_points_build) used for moving
meshes: on lattice points it reproduces the lattice builder's mesh; on
moved points the cells tile the box, faces are perpendicular bisectors and
every wall face is the mirror face of its cell."""
import numpy as np
import pytest

from rich_amd import mesh as rmesh

BOX = (-1.0, 1.0, -1.0, 1.0)


def _pairs(m):
    o = m.orig(np.stack([m.face_first, m.face_second]))
    return {tuple(sorted(p)) + (int(min(a, b) >= m.n_cells or max(a, b) >= m.n_cells),)
            for p, a, b in zip(o.T, m.face_first, m.face_second)}


def test_points_reproduce_lattice_mesh():
    m0 = rmesh.lattice(20, 14, BOX, jitter=1e-2, seed=4)
    N = m0.n_cells
    m1 = rmesh.points(m0.pt_x[:N], m0.pt_y[:N], BOX)
    assert (m1.n_cells, m1.n_faces, m1.n_ghosts) == (m0.n_cells, m0.n_faces, m0.n_ghosts)
    assert _pairs(m1) == _pairs(m0)
    np.testing.assert_allclose(m1.volume, m0.volume, rtol=1e-13)
    np.testing.assert_allclose(m1.tess_volume, m0.tess_volume, rtol=1e-13)
    np.testing.assert_allclose(m1.cm_x[:N], m0.cm_x[:N], atol=1e-14)
    np.testing.assert_allclose(m1.cm_y[:N], m0.cm_y[:N], atol=1e-14)


@pytest.mark.parametrize("shake", [0.1, 0.35])
def test_moved_points_mesh(shake):
    m0 = rmesh.lattice(24, 24, BOX, jitter=1e-3, seed=9)
    N = m0.n_cells
    rng = np.random.default_rng(2)
    dx = 2.0 / 24
    x = np.clip(m0.pt_x[:N] + shake * dx * rng.uniform(-1, 1, N), -1 + 1e-3, 1 - 1e-3)
    y = np.clip(m0.pt_y[:N] + shake * dx * rng.uniform(-1, 1, N), -1 + 1e-3, 1 - 1e-3)
    m = rmesh.points(x, y, BOX)
    assert m.n_cells == N
    assert abs(m.tess_volume.sum() - 4.0) < 1e-12
    assert abs(m.volume.sum() - 4.0) < 1e-12
    assert np.all(np.abs(m.cm_x[:N]) < 1) and np.all(np.abs(m.cm_y[:N]) < 1)
    # faces are perpendicular bisectors of their two points
    a, b = m.face_first, m.face_second
    for vx, vy in ((m.v1x, m.v1y), (m.v2x, m.v2y)):
        da = np.hypot(vx - m.pt_x[a], vy - m.pt_y[a])
        db = np.hypot(vx - m.pt_x[b], vy - m.pt_y[b])
        np.testing.assert_allclose(da, db, rtol=1e-9, atol=1e-12)
    # every ghost borders only the cell it mirrors, on the wall
    g = np.where(a >= N, a, np.where(b >= N, b, -1))
    r = np.where(a >= N, b, a)
    wall = g >= 0
    assert np.all(m.ghost_orig[g[wall] - N] == r[wall])
    mid_x = 0.5 * (m.v1x[wall] + m.v2x[wall])
    mid_y = 0.5 * (m.v1y[wall] + m.v2y[wall])
    on_wall = np.minimum(np.minimum(np.abs(mid_x - 1), np.abs(mid_x + 1)),
                         np.minimum(np.abs(mid_y - 1), np.abs(mid_y + 1)))
    assert np.all(on_wall < 1e-12)
    # CSR: ascending faces, each real side listed
    for i in range(0, N, 37):
        fs = m.csr_face[m.csr_off[i]:m.csr_off[i + 1]]
        assert np.all(np.diff(fs) > 0)
        assert np.all((a[fs] == i) | (b[fs] == i))
