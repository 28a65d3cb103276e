"""The Hilbert re-sort of the moving-mesh run (hdsim2d.cpp:389-399: every
25 cycles the points are reordered and the cells follow them).

* `hilbert_order` walks a 2^k x 2^k lattice in unit steps, starting in a
  corner and ending in the next corner of the same side (the curve's defining
  property), and is a permutation for any points;
* a run that re-sorts its cells is the same run as one that does not: the
  points, time steps and cells agree once put back in one order, up to the
  rounding of sums whose terms come in another order."""
import numpy as np
import pytest

import cases
from oracle_drivers import oracle_moving_mesh
from rich_amd.moving import hilbert_order


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_lattice_walk_is_a_hilbert_curve(k):
    n = 1 << k
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x, y = (i.ravel() + 0.5) / n, (j.ravel() + 0.5) / n
    o = hilbert_order(x, y, (0.0, 1.0, 0.0, 1.0), bits=k)
    assert np.array_equal(np.sort(o), np.arange(n * n))
    steps = np.abs(np.diff(i.ravel()[o])) + np.abs(np.diff(j.ravel()[o]))
    assert np.all(steps == 1)
    first, last = (i.ravel()[o[0]], j.ravel()[o[0]]), (i.ravel()[o[-1]], j.ravel()[o[-1]])
    assert first == (0, 0) and last in ((n - 1, 0), (0, n - 1))


def test_random_points_permutation():
    rng = np.random.default_rng(4)
    x, y = rng.uniform(-2, 3, 5000), rng.uniform(-1, 1, 5000)
    o = hilbert_order(x, y, (-2.0, 3.0, -1.0, 1.0))
    assert np.array_equal(np.sort(o), np.arange(5000))
    # neighbours along the curve are close: the mean step is near the
    # spacing of the points, far below that of a random order
    d = np.hypot(np.diff(x[o]), np.diff(y[o])).mean()
    assert d < 0.1 * np.hypot(np.diff(x), np.diff(y)).mean()


def _canonical(r):
    o = np.lexsort((r.y, r.x))
    return r.x[o], r.y[o], r.cells[:, o]


def test_resorted_run_is_the_same_run():
    m, c = cases.smooth_box_case()
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    plain = oracle_moving_mesh(m.pt_x[:N], m.pt_y[:N], c, box, resort_every=0)
    sorted_ = oracle_moving_mesh(m.pt_x[:N], m.pt_y[:N], c, box, resort_every=3)
    for step in range(12):
        dt0, dt1 = plain.step(), sorted_.step()
        assert abs(dt0 - dt1) <= 1e-13 * dt0, step
    assert sorted_.reorders > 0 and plain.reorders == 0  # the re-sort changed the order
    x0, y0, c0 = _canonical(plain)
    x1, y1, c1 = _canonical(sorted_)
    np.testing.assert_allclose(x1, x0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(y1, y0, rtol=0, atol=1e-14)
    scale = np.abs(c0).max(axis=1, keepdims=True)
    assert np.all(np.abs(c1 - c0) <= 1e-13 * scale)
