"""The reference's riemann_profiles_messy_grid test (tests/newtonian/
two_dimensional/riemann_profiles_messy_grid/test.cpp, test.py) on the
oracle: a 30 x 30 Cartesian grid in the unit box with rigid walls,
Lagrangian + RoundCells points (chi 0.15, eta 0.02; the overshoot fix of
RoundCells' default outer boundary, the box [-1, 1]^2, round_cells.cpp:10-11,
131-132), StationaryBox edges, HLLC through SimpleFluxCalculator,
TimeAdvance, CFL 0.3, pressure 2 left of x = 0.5 and 1 right of it, run
while t < 0.225.  The grader's goodness of fit against the exact Riemann
solution, sqrt(sum (sim - exact)^2 / (max sim - min sim) / n) over the
cells sorted by their points' x, must stay below 0.1 (density), 0.11
(pressure) and 0.11 (x velocity)."""
import numpy as np

import exact_riemann
from oracle_drivers import OracleBackend, OracleMotion
from rich_amd import mesh as rmesh
from rich_amd.moving import MovingMesh

GAMMA = 5.0 / 3.0
REFERENCE_FIX_BOX = (-1.0, 1.0, -1.0, 1.0)


class OracleMotionReferenceFix(OracleMotion):
    def point_fix(self, mesh, vx, vy, dt, box):
        return super().point_fix(mesh, vx, vy, dt, REFERENCE_FIX_BOX)


def messy_grid_run(backend, motion):
    m = rmesh.lattice(30, 30, (0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=1, bc=rmesh.RIGID)
    N = m.n_cells
    x = m.pt_x[:N]
    cells = np.vstack([np.ones(N), np.where(x < 0.5, 2.0, 1.0), np.zeros(N), np.zeros(N)])
    return MovingMesh(m.pt_x[:N], m.pt_y[:N], cells, (0.0, 1.0, 0.0, 1.0), backend,
                      motion=motion, integrator=1)


def goodness_of_fit(sim, exact):
    sim = np.asarray(sim)
    return float(np.sqrt(np.sum((sim - exact) ** 2) / (sim.max() - sim.min()) / len(sim)))


def grade(run):
    return grade_state(run.x, run.time, run.cells)


def grade_state(x, t, cells):
    ids = np.argsort(x)
    xs = x[ids]
    d, u, p = exact_riemann.sample(xs, t, 0.5, (1.0, 0.0, 2.0), (1.0, 0.0, 1.0), GAMMA)
    return (goodness_of_fit(cells[0][ids], d), goodness_of_fit(cells[1][ids], p),
            goodness_of_fit(cells[2][ids], u))


def test_riemann_messy_grid_oracle():
    run = messy_grid_run(OracleBackend(order=1), OracleMotionReferenceFix())
    while run.time < 0.225:
        run.step()
    gd, gp, gv = grade(run)
    print(f"t = {run.time:.4f}: goodness of fit {gd:.4f} {gp:.4f} {gv:.4f}")
    assert gd < 0.1 and gp < 0.11 and gv < 0.11
    # the points moved with the flow: distance to the nearest grid point
    gx = (np.floor(run.x * 30) + 0.5) / 30
    gy = (np.floor(run.y * 30) + 0.5) / 30
    assert np.hypot(run.x - gx, run.y - gy).max() > 0.005


def test_riemann_messy_grid_second_order_oracle():
    """The reference's riemann_profiles_messy_grid_2 run: the same problem
    at second order (LinearGaussImproved with RigidWallGenerator,
    ModularFluxCalculator, TimeAdvance2Heun); bounds 0.067, 0.053 and
    0.062."""
    m = rmesh.lattice(30, 30, (0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=1, bc=rmesh.RIGID)
    N = m.n_cells
    x = m.pt_x[:N]
    cells = np.vstack([np.ones(N), np.where(x < 0.5, 2.0, 1.0), np.zeros(N), np.zeros(N)])
    run = MovingMesh(m.pt_x[:N], m.pt_y[:N], cells, (0.0, 1.0, 0.0, 1.0), OracleBackend(order=2),
                     motion=OracleMotionReferenceFix(), integrator=2)
    while run.time < 0.225:
        run.step()
    gd, gp, gv = grade(run)
    print(f"t = {run.time:.4f}: goodness of fit {gd:.4f} {gp:.4f} {gv:.4f}")
    assert gd < 0.067 and gp < 0.053 and gv < 0.062
