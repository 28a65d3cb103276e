"""The reference's sedov_taylor_messy_grid test (tests/newtonian/
two_dimensional/sedov_taylor_messy_grid/test.cpp, test.py) on the oracle:
a 50 x 50 Cartesian grid in the unit box with rigid walls, Lagrangian
points, StationaryBox edges, HLLC through SimpleFluxCalculator,
TimeAdvance, CFL 0.3, density 1, pressure 1e4 for points within 0.06 of the
corner (0, 0) and 0.01 elsewhere, 100 steps.

The grader finds the shock from the pressure maximum (a parabola through
the five points around it, by distance from the corner), scales the
Sedov-Taylor profiles (cylindrical, gamma 5/3) by the simulation's values
there, and requires mean absolute differences below 0.5 (density), 14
(pressure) and 1.5 (radial velocity).  The closed form is pinned by the
textbook's energy constants."""
import numpy as np
import pytest

import sedov_taylor
from oracle_drivers import OracleBackend, OracleLagrangianMotion
from rich_amd import mesh as rmesh
from rich_amd.moving import MovingMesh

GAMMA = 5.0 / 3.0


@pytest.mark.parametrize("g,j,xi0", [(1.4, 3, 1.0328), (5.0 / 3.0, 3, 1.1517),
                                     (1.4, 2, 1.0040)])
def test_closed_form_energy_constants(g, j, xi0):
    """R = xi0 (E t^2 / rho0)^(1/(j+2)): xi0 = 1.0328 and 1.1517 for a
    spherical blast in gamma 1.4 and 5/3 gases, 1.0040 for a cylindrical
    one in gamma 1.4 (Sedov's values), and every profile ratio is 1 at the
    shock."""
    alpha = sedov_taylor.energy_constant(g, j)
    assert abs(alpha ** (-1.0 / (j + 2)) - xi0) < 2e-4
    lam, f, gd, h = sedov_taylor.profiles(sedov_taylor.v_range(g, j)[1], g, j)
    np.testing.assert_allclose([lam, f, gd, h], 1.0, rtol=1e-12)


def sedov_grid_run(backend, motion):
    m = rmesh.lattice(50, 50, (0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=1, bc=rmesh.RIGID)
    N = m.n_cells
    cells = sedov_grid_cells(m)
    return MovingMesh(m.pt_x[:N], m.pt_y[:N], cells, (0.0, 1.0, 0.0, 1.0), backend,
                      motion=motion, integrator=1)


def grade(run):
    """The grader's mean absolute differences (density, pressure, velocity)."""
    return grade_cells(run.x, run.y, run.cells)


def grade_cells(x, y, cells):
    radius = np.hypot(x, y)
    d, p = cells[0], cells[1]
    v = (x * cells[2] + y * cells[3]) / radius
    # each field sorted by (radius, value), as the grader's sorted(zip(...))
    rs = np.sort(radius)
    ds, ps, vs = (a[np.lexsort((a, radius))] for a in (d, p, v))
    mi = int(np.argmax(ps))
    fit = np.polyfit(rs[mi - 2:mi + 3], ps[mi - 2:mi + 3], 2)
    front = -0.5 * fit[1] / fit[0]
    fd, fp, fv = (np.interp(front, rs, a) for a in (ds, ps, vs))
    # the grader's table: 1000 values of the similarity variable spaced
    # evenly from 1e-6 past the centre (its V is twice this one for j = 2)
    lam, f, gd, h = sedov_taylor.tables(GAMMA, 2, n=1000, offset=0.5e-6)
    rad = front * lam

    def analytic(r, shape, at_front, upstream):
        return np.where(r > front, upstream, np.interp(r, rad, at_front * shape))

    l1 = [float(np.mean(np.abs(analytic(radius, shape, at, up) - num)))
          for shape, at, up, num in ((gd, fd, 1.0, d), (h, fp, 0.01, p), (f, fv, 0.0, v))]
    return front, l1


def test_sedov_taylor_grid_oracle():
    run = sedov_grid_run(OracleBackend(order=1), OracleLagrangianMotion())
    for _ in range(100):
        run.step()
    front, (ld, lp, lv) = grade(run)
    print(f"t = {run.time:.4e}, shock at r = {front:.4f}: L1 {ld:.4f} {lp:.4f} {lv:.4f}")
    assert 0.1 < front < 0.9
    assert ld < 0.5 and lp < 14 and lv < 1.5


def sedov_grid_cells(m):
    N = m.n_cells
    r = np.hypot(m.pt_x[:N], m.pt_y[:N])
    return np.vstack([np.ones(N), np.where(r < 0.06, 1e4, 0.01), np.zeros(N), np.zeros(N)])


def test_eulerian_sedov_taylor_oracle():
    """The reference's eulerian_sedov_taylor run: the same grid and blast on
    a static mesh, 100 TimeAdvance steps of SimpleFluxCalculator; the same
    grader with the pressure bound 15."""
    import pyoracle as po
    m = rmesh.lattice(50, 50, (0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=1, bc=rmesh.RIGID)
    N = m.n_cells
    o = po.Oracle(m, order=1)
    c = sedov_grid_cells(m)
    c, _, dts = o.run(c, o.init_extensives(c), 100, integrator=1)
    front, (ld, lp, lv) = grade_cells(m.pt_x[:N], m.pt_y[:N], c)
    print(f"t = {dts.sum():.4e}, shock at r = {front:.4f}: L1 {ld:.4f} {lp:.4f} {lv:.4f}")
    assert 0.1 < front < 0.9
    assert ld < 0.5 and lp < 15 and lv < 1.5
