"""Pins the oracle's moving-mesh inputs against properties the geometry
fixes, not against their formulas:

* RoundCells::calc_dw (round_cells.cpp:63-76): the correction w - v points
  from the mesh point to the cell's CM, has length chi max(c, |v|) d / R,
  and vanishes where d < 0.9 eta R;
* CorrectPointsOverShoot (round_cells.cpp:14-61): no corrected point
  leaves the box in one step, a point within 0.1 R of a wall it heads for
  turns back, and points far from the walls keep their velocity;
* CalcFaceVelocity (VoronoiMesh.cpp:146-151) under StationaryBox
  (stationary_box.cpp:20-40): a uniform motion moves every interior face
  with it, mirror faces stay still, and the normal component is the speed
  at which the perpendicular bisector of the two moving points sweeps over
  the face centroid (measured by moving the points a little each way).
"""
import numpy as np
import pytest

import cases
import pyoracle as po
from rich_amd import mesh as rmesh

G = 5.0 / 3.0


def _mesh():
    return rmesh.lattice(24, 20, (-1.0, 1.0, -0.8, 0.8), jitter=5e-2, seed=17, bc=rmesh.RIGID)


def _cells(m, seed=1):
    c = cases.rough_cells(m, seed=seed)
    c[0] = np.abs(c[0]) + 0.3
    c[1] = np.abs(c[1]) + 0.3
    c[2] *= 3.0  # some cells move faster than their sound speed
    return c


def test_round_cells_correction():
    m = _mesh()
    N = m.n_cells
    cells = _cells(m)
    r = np.vstack([m.pt_x[:N], m.pt_y[:N]])
    s = np.vstack([m.cm_x[:N], m.cm_y[:N]]) - r
    d = np.hypot(*s)
    R = np.sqrt(m.tess_volume / np.pi)
    # eta that puts the 0.9 eta R threshold in the middle of the d/R range
    eta = np.median(d / R) / 0.9
    chi = 0.15
    w = po.Oracle(m).point_velocity(cells, chi=chi, eta=eta)
    dw = w - cells[2:]
    on = d >= 0.9 * eta * R
    assert 0.2 * N < on.sum() < 0.8 * N
    assert not dw[:, ~on].any()
    c = np.sqrt(G * cells[1] / cells[0])
    speed = np.maximum(c, np.hypot(cells[2], cells[3]))
    assert (np.hypot(cells[2], cells[3]) > c).sum() > 10
    np.testing.assert_allclose(np.hypot(*dw[:, on]), chi * speed[on] * d[on] / R[on], rtol=1e-13)
    cross = dw[0, on] * s[1, on] - dw[1, on] * s[0, on]
    assert np.all(np.abs(cross) <= 1e-13 * np.hypot(*dw[:, on]) * d[on])
    assert np.all(dw[0, on] * s[0, on] + dw[1, on] * s[1, on] > 0)


def test_overshoot_fix_keeps_points_in_the_box():
    m = _mesh()
    N = m.n_cells
    box = (-1.0, 1.0, -0.8, 0.8)
    rng = np.random.default_rng(3)
    w = rng.standard_normal((2, N)) * 2.0
    # every wall cell heads for its wall
    px, py = m.pt_x[:N], m.pt_y[:N]
    w[0] += np.where(px > 0.9, 3.0, 0.0) - np.where(px < -0.9, 3.0, 0.0)
    w[1] += np.where(py > 0.7, 3.0, 0.0) - np.where(py < -0.7, 3.0, 0.0)
    dt = 0.02
    fixed = po.Oracle(m).point_fix(w, dt, box)
    nx, ny = px + dt * fixed[0], py + dt * fixed[1]
    assert np.all((nx > box[0]) & (nx < box[1]) & (ny > box[2]) & (ny < box[3]))
    crossing = ((px + 2 * dt * w[0] > box[1]) | (px + 2 * dt * w[0] < box[0]) |
                (py + 2 * dt * w[1] > box[3]) | (py + 2 * dt * w[1] < box[2]))
    assert crossing.sum() > 20
    # a point whose velocity is not changed by the first test is untouched
    # by every later one unless it would cross within two steps
    far = ~crossing
    np.testing.assert_array_equal(fixed[:, far], w[:, far])
    assert np.all(fixed[:, crossing] != w[:, crossing])
    # within 0.1 R of the right wall and heading for it: turned back
    m2 = _near_wall_mesh()
    N2 = m2.n_cells
    R2 = np.sqrt(m2.tess_volume / np.pi)
    w2 = np.zeros((2, N2))
    w2[0] = 1.0
    f2 = po.Oracle(m2).point_fix(w2, dt, (0.0, 1.0, 0.0, 1.0))
    close = (1.0 - m2.pt_x[:N2] < 0.1 * R2)
    assert close.any()
    assert np.all(f2[0, close] < 0)


def _near_wall_mesh():
    """Points of a lattice with the last column pressed against the right
    wall (within 0.1 R of it)."""
    n = 10
    xs = (np.arange(n) + 0.5) / n
    x, y = np.meshgrid(xs, xs, indexing="ij")
    x = x.ravel().copy()
    y = y.ravel().copy()
    x[x > 0.9] = 1.0 - 0.002
    return rmesh.points(x, y, (0.0, 1.0, 0.0, 1.0))


def test_uniform_motion_moves_faces_with_it():
    for bc in (rmesh.RIGID, rmesh.PERIODIC):
        m = rmesh.lattice(16, 12, (0.0, 1.6, 0.0, 1.2), jitter=5e-2, seed=2, bc=bc)
        N = m.n_cells
        w = np.vstack([np.full(N, 0.37), np.full(N, -1.25)])
        fw = po.Oracle(m).face_velocity(w)
        mirror = m.orig(m.face_first) == m.orig(m.face_second)
        if bc == rmesh.RIGID:
            assert mirror.sum() > 20
            assert not fw[:, mirror].any()
        else:
            assert not mirror.any()
        np.testing.assert_array_equal(fw[0, ~mirror], 0.37)
        np.testing.assert_array_equal(fw[1, ~mirror], -1.25)


@pytest.mark.parametrize("bc", [rmesh.RIGID, rmesh.PERIODIC])
def test_face_normal_speed_is_the_bisector_speed(bc):
    m = rmesh.lattice(16, 12, (0.0, 1.6, 0.0, 1.2), jitter=5e-2, seed=2, bc=bc)
    N = m.n_cells
    w = np.random.default_rng(4).standard_normal((2, N))
    fw = po.Oracle(m).face_velocity(w)
    a, b = m.face_first, m.face_second
    oa, ob = m.orig(a), m.orig(b)
    mirror = oa == ob
    rl = np.vstack([m.pt_x[a], m.pt_y[a]])
    rr = np.vstack([m.pt_x[b], m.pt_y[b]])
    wl, wr = w[:, oa], w[:, ob]
    f = np.vstack([(m.v1x + m.v2x) * 0.5, (m.v1y + m.v2y) * 0.5])

    def offset(tau):
        """Signed distance of f from the bisector of the moved points,
        along the unit normal rl -> rr."""
        pl, pr = rl + tau * wl, rr + tau * wr
        d = pr - pl
        n = d / np.hypot(*d)
        mid = 0.5 * (pl + pr)
        return np.sum((f - mid) * n, axis=0)

    tau = 1e-6
    # the face sweeps over f at -d(offset)/dt
    sweep = -(offset(tau) - offset(-tau)) / (2 * tau)
    n = (rr - rl) / np.hypot(*(rr - rl))
    normal = np.sum(fw * n, axis=0)
    sel = ~mirror
    np.testing.assert_allclose(normal[sel], sweep[sel], rtol=1e-7, atol=1e-9)
    # the tangential component is the two points' mean
    t = np.vstack([-n[1], n[0]])
    np.testing.assert_allclose(np.sum(fw * t, axis=0)[sel],
                               np.sum(0.5 * (wl + wr) * t, axis=0)[sel], rtol=1e-12, atol=1e-12)
