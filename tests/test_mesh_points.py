"""The point-set mesh builders used for moving meshes (rich_amd.mesh.points:
rm_box_points_build, i.e. rm_delaunay of the points and their wall
reflections, then the faces; BoxMesher: the previous triangulation repaired
by flips): on lattice points they reproduce the lattice builder's mesh; on
moved points the cells tile the box, faces are perpendicular bisectors and
every wall face is the mirror face of its cell.  rm_delaunay is checked
against Qhull and against the empty-circumcircle property by brute force."""
import ctypes
import numpy as np
import pytest

from rich_amd import mesh as rmesh

BOX = (-1.0, 1.0, -1.0, 1.0)


def qhull_points_mesh(x, y, box, margin):
    """The same ghosts as rich_amd.mesh.points, triangulated by Qhull
    (scipy.spatial.Delaunay) instead, then rm_points_build: an independent
    triangulation to compare the native one with."""
    from scipy.spatial import Delaunay
    N = len(x)
    x0, x1, y0, y1 = box
    gx, gy, go, gw = [], [], [], []
    for wall, near, fx, fy in (
            (0, x1 - x < margin, lambda: x + 2 * (x1 - x), lambda: y + 0.0),
            (1, y1 - y < margin, lambda: x + 0.0, lambda: y + 2 * (y1 - y)),
            (2, x - x0 < margin, lambda: x + 2 * (x0 - x), lambda: y + 0.0),
            (3, y - y0 < margin, lambda: x + 0.0, lambda: y + 2 * (y0 - y))):
        idx = np.nonzero(near)[0]
        gx.append(fx()[idx])
        gy.append(fy()[idx])
        go.append(idx)
        gw.append(np.full(len(idx), wall))
    X = np.concatenate([x] + gx)
    Y = np.concatenate([y] + gy)
    gorig, gwall = np.concatenate(go), np.concatenate(gw)
    tri = Delaunay(np.column_stack([X, Y])).simplices.astype(np.int64)
    tri = tri[(tri < N).any(axis=1)]  # triangles with a real vertex
    used = np.unique(tri[tri >= N]) - N
    new_id = np.full(len(gorig), -1, np.int64)
    new_id[used] = N + np.arange(len(used))
    tri = np.where(tri >= N, new_id[np.maximum(tri - N, 0)], tri)
    return rmesh.from_triangles(np.concatenate([x, X[N:][used]]),
                                np.concatenate([y, Y[N:][used]]), gorig[used], gwall[used],
                                tri, box)


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


def _tri_set(t):
    return {tuple(sorted(r)) for r in np.asarray(t).tolist()}


def _empty_circles(x, y, tri, rel=1e-9):
    """Brute force: no point lies inside any triangle's circumcircle (beyond
    round-off), and every triangle is counter-clockwise."""
    a, b, c = (np.stack([x[tri[:, k]], y[tri[:, k]]], 1) for k in range(3))
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    assert np.all(cross > 0)
    B, C = b - a, c - a
    d = 2 * (B[:, 0] * C[:, 1] - B[:, 1] * C[:, 0])
    b2, c2 = (B * B).sum(1), (C * C).sum(1)
    ux = (C[:, 1] * b2 - B[:, 1] * c2) / d
    uy = (B[:, 0] * c2 - C[:, 0] * b2) / d
    cx, cy, r2 = a[:, 0] + ux, a[:, 1] + uy, ux * ux + uy * uy
    for i in range(len(x)):
        d2 = (x[i] - cx) ** 2 + (y[i] - cy) ** 2
        assert np.all(d2 >= r2 * (1 - rel)), i


def test_delaunay_matches_qhull_in_general_position():
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(11)
    for n in (3, 4, 10, 300, 5000):
        x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        t = rmesh.delaunay(x, y)
        assert _tri_set(t) == _tri_set(Delaunay(np.column_stack([x, y])).simplices)


def test_delaunay_brute_force_and_degenerate_inputs():
    rng = np.random.default_rng(5)
    x, y = rng.uniform(0, 1, 400), rng.uniform(0, 1, 400)
    _empty_circles(x, y, rmesh.delaunay(x, y))
    # a square lattice: every quad cocircular, any diagonal is Delaunay
    gx, gy = np.meshgrid(np.arange(12.0), np.arange(9.0))
    x, y = gx.ravel(), gy.ravel()
    t = rmesh.delaunay(x, y)
    assert len(t) == 2 * 11 * 8
    _empty_circles(x, y, t)
    # points on a circle and collinear runs
    a = np.linspace(0, 2 * np.pi, 40, endpoint=False)
    x = np.concatenate([np.cos(a), np.linspace(-0.5, 0.5, 11), [0.0]])
    y = np.concatenate([np.sin(a), np.zeros(11), [0.3]])
    _empty_circles(x, y, rmesh.delaunay(x, y))


def test_delaunay_rejects_duplicates():
    lib = rmesh._lib()
    x = np.array([0.0, 1.0, 0.0, 1.0, 1.0])
    y = np.array([0.0, 0.0, 1.0, 1.0, 0.0])
    tri = np.empty((10, 3), np.int32)
    nt = ctypes.c_int32()
    rc = lib.rm_delaunay(5, x.ctypes.data, y.ctypes.data, tri.ctypes.data, 10, ctypes.byref(nt))
    assert rc < 0
    with pytest.raises(ValueError):
        rmesh.delaunay(x, y)
    with pytest.raises(ValueError):
        rmesh.points(np.concatenate([x, [0.5]]), np.concatenate([y, [0.5]]), (-0.5, 1.5, -0.5, 1.5))


def _same_mesh(a, b):
    pa = {tuple(sorted(p)) for p in zip(a.face_first, a.face_second)}
    pb = {tuple(sorted(p)) for p in zip(b.face_first, b.face_second)}
    assert (a.n_faces, a.n_ghosts) == (b.n_faces, b.n_ghosts)
    assert pa == pb
    np.testing.assert_allclose(a.volume, b.volume, rtol=1e-12)
    np.testing.assert_allclose(a.tess_volume, b.tess_volume, rtol=1e-12)
    np.testing.assert_allclose(a.cm_x, b.cm_x, rtol=0, atol=1e-13)
    np.testing.assert_allclose(a.cm_y, b.cm_y, rtol=0, atol=1e-13)


def test_points_equals_qhull_path():
    m0 = rmesh.lattice(30, 30, BOX, jitter=1e-3, seed=2)
    N = m0.n_cells
    rng = np.random.default_rng(8)
    x = np.clip(m0.pt_x[:N] + 0.2 * (2 / 30) * rng.uniform(-1, 1, N), -0.999, 0.999)
    y = np.clip(m0.pt_y[:N] + 0.2 * (2 / 30) * rng.uniform(-1, 1, N), -0.999, 0.999)
    a = rmesh.points(x, y, BOX)
    b = qhull_points_mesh(x, y, BOX, 3.0 * np.sqrt(4.0 / N))
    _same_mesh(a, b)


def test_box_mesher_repairs_match_fresh_builds():
    """Small steps are repaired by flips, a large jump triggers a fresh
    triangulation; every mesh equals the fresh one of the same points."""
    m0 = rmesh.lattice(40, 40, BOX, jitter=1e-3, seed=6)
    N = m0.n_cells
    x, y = m0.pt_x[:N].copy(), m0.pt_y[:N].copy()
    B = rmesh.BoxMesher(N, BOX)
    rng = np.random.default_rng(9)
    h = 2.0 / 40
    for step in range(12):
        _same_mesh(B.mesh(x, y), rmesh.points(x, y, BOX))
        amp = 0.6 if step == 7 else 0.03  # step 8 scrambles the points
        x = np.clip(x + amp * h * rng.standard_normal(N), -0.999, 0.999)
        y = np.clip(y + amp * h * rng.standard_normal(N), -0.999, 0.999)
    builds, repairs = B.counts()
    assert builds >= 2 and repairs >= 5, (builds, repairs)
    B.close()


_SEQ = """
import sys, numpy as np
sys.path.insert(0, {pkg!r})
from rich_amd import mesh as rmesh
m0 = rmesh.lattice(300, 300, (-1.0, 1.0, -1.0, 1.0), jitter=1e-3, seed=12)
N = m0.n_cells
x, y = m0.pt_x[:N].copy(), m0.pt_y[:N].copy()
B = rmesh.BoxMesher(N, (-1.0, 1.0, -1.0, 1.0))
rng = np.random.default_rng(1)
out = {{}}
for s in range(3):
    m = B.mesh(x, y)
    for k in ("face_first", "face_second", "v1x", "v2y", "csr_face", "volume", "cm_x"):
        out[f"{{k}}_{{s}}"] = getattr(m, k)
    x = np.clip(x + 0.05 * (2 / 300) * rng.standard_normal(N), -0.999, 0.999)
    y = np.clip(y + 0.05 * (2 / 300) * rng.standard_normal(N), -0.999, 0.999)
np.savez({path!r}, **out)
"""


def test_mesher_threads_do_not_change_the_mesh(tmp_path):
    """The mesher's host threads split work in fixed chunks joined in order:
    one thread and many give the same meshes, bit for bit."""
    import os
    import subprocess
    import sys
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "huji-rich_amd")
    res = {}
    for t in ("1", "8"):
        path = str(tmp_path / f"m{t}.npz")
        env = dict(os.environ, RICH_MESH_THREADS=t)
        subprocess.run([sys.executable, "-c", _SEQ.format(pkg=pkg, path=path)], env=env,
                       check=True)
        res[t] = np.load(path)
    for k in res["1"].files:
        np.testing.assert_array_equal(res["1"][k], res["8"][k], err_msg=k)


def _same_geometry(a, b):
    """The same Voronoi mesh up to the numbering of ghosts: faces between
    real cells, each real cell's volumes and CM, and the ghosts as
    (original, position)."""
    N = a.n_cells
    assert (b.n_cells, b.n_faces, b.n_ghosts) == (N, a.n_faces, a.n_ghosts)

    def real_pairs(m):
        return {tuple(sorted(p)) for p in zip(m.face_first, m.face_second) if max(p) < N}

    def ghosts(m):
        return sorted((int(o), round(float(x), 12), round(float(y), 12)) for o, x, y in
                      zip(m.ghost_orig, m.pt_x[N:], m.pt_y[N:]))

    assert real_pairs(a) == real_pairs(b)
    assert ghosts(a) == ghosts(b)
    np.testing.assert_allclose(a.volume, b.volume, rtol=1e-12)
    np.testing.assert_allclose(a.tess_volume, b.tess_volume, rtol=1e-12)
    np.testing.assert_allclose(a.cm_x[:N], b.cm_x[:N], rtol=0, atol=1e-13)
    np.testing.assert_allclose(a.cm_y[:N], b.cm_y[:N], rtol=0, atol=1e-13)


def test_box_mesher_follows_a_renumbering():
    """rm_mover_permute: after the points are renumbered the kept
    triangulation is repaired (no fresh build) into the mesh a fresh build
    of the renumbered points gives; anything but a permutation is refused."""
    m0 = rmesh.lattice(30, 30, BOX, jitter=1e-3, seed=4)
    N = m0.n_cells
    rng = np.random.default_rng(3)
    x, y = m0.pt_x[:N].copy(), m0.pt_y[:N].copy()
    B = rmesh.BoxMesher(N, BOX)
    B.mesh(x, y)
    order = rng.permutation(N)
    B.permute(order)
    x, y = x[order], y[order]
    x = np.clip(x + 0.03 * (2 / 30) * rng.standard_normal(N), -0.999, 0.999)
    y = np.clip(y + 0.03 * (2 / 30) * rng.standard_normal(N), -0.999, 0.999)
    _same_geometry(B.mesh(x, y), rmesh.points(x, y, BOX))
    assert B.counts() == (1, 1)
    bad = order.copy()
    bad[0] = bad[1]
    with pytest.raises(ValueError):
        B.permute(bad)
    B.close()
