"""The reference's simple_waves_plm_2 test (tests/newtonian/
two_dimensional/simple_waves_plm_2/test.cpp, test.py) on the oracle: a
30 x 30 Cartesian grid in the unit box with rigid walls, LinearGaussImproved
with RigidWallGenerator, HLLC through ModularFluxCalculator, Lagrangian +
RoundCells points (overshoot fix against RoundCells' default box
[-1, 1]^2), StationaryBox edges, TimeAdvance2Heun, CFL 0.3, gamma
1.66666666667, run while t < 0.02.

The initial state is a simple wave: density 1 + 10 ((((x - 0.5) / 0.3)^2 - 1)^4
within 0.3 of x = 0.5 (Collela, collela.cpp), pressure = density^gamma,
and the velocity that keeps the Riemann invariant v - 2 c / (gamma - 1)
at its edge value (simple_waves_ideal_gas.cpp), at each cell's CM.  The
grader moves each initial point by t (c + v) and compares the final cells,
interpolated there, with the initial values: goodness of fit below 0.09
(density), 0.26 (pressure) and 0.11 (velocity)."""
import numpy as np

from oracle_drivers import OracleBackend, OracleMotion
from rich_amd import mesh as rmesh
from rich_amd.moving import MovingMesh

GAMMA = 1.66666666667
FIX_BOX = (-1.0, 1.0, -1.0, 1.0)


class OracleMotionDefaultFix(OracleMotion):
    def __init__(self):
        super().__init__(gamma=GAMMA)

    def point_fix(self, mesh, vx, vy, dt, box):
        return super().point_fix(mesh, vx, vy, dt, FIX_BOX)


def simple_wave_cells(x):
    xc = x - 0.5
    d = np.where(np.abs(xc) > 0.3, 1.0, 1.0 * (1 + 10 * ((xc / 0.3) ** 2 - 1) ** 4))
    p = 1.0 * d ** GAMMA
    c = np.sqrt(GAMMA * p / d)
    c_edge = np.sqrt(GAMMA * 1.0 / 1.0)
    rv_edge = 0 - 2 * c_edge / (GAMMA - 1)
    v = rv_edge - (0 - 2 * c / (GAMMA - 1))
    return np.vstack([d, p, v, np.zeros_like(x)])


def simple_wave_mesh():
    m = rmesh.lattice(30, 30, (0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=1, bc=rmesh.RIGID)
    N = m.n_cells
    return m, simple_wave_cells(m.cm_x[:N])


def goodness_of_fit(a1, a2):
    a1 = np.asarray(a1)
    return float(np.sqrt(np.sum((a1 - np.asarray(a2)) ** 2) / (a1.max() - a1.min()) / len(a1)))


def grade(x0, c0, x, c, t):
    ids0 = np.argsort(x0, kind="stable")
    xs0, ds0, ps0, vs0 = x0[ids0], c0[0][ids0], c0[1][ids0], c0[2][ids0]
    cs0 = np.sqrt(GAMMA * ps0 / ds0)
    ids = np.argsort(x, kind="stable")
    xs, ds, ps, vs = x[ids], c[0][ids], c[1][ids], c[2][ids]
    xa = xs0 + t * (cs0 + vs0)
    inside = (xa < xs.max()) & (xa > xs.min())
    xi = xa[inside]
    return (goodness_of_fit(np.interp(xi, xs, ds), ds0[inside]),
            goodness_of_fit(np.interp(xi, xs, ps), ps0[inside]),
            goodness_of_fit(np.interp(xi, xs, vs), vs0[inside]))


def test_simple_waves_oracle():
    m, c = simple_wave_mesh()
    N = m.n_cells
    run = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, (0.0, 1.0, 0.0, 1.0),
                     OracleBackend(order=2, gamma=GAMMA), motion=OracleMotionDefaultFix())
    while run.time < 0.02:
        run.step()
    g = grade(m.pt_x[:N], c, run.x, run.cells, run.time)
    print(f"t = {run.time:.4f}: goodness of fit {g[0]:.4f} {g[1]:.4f} {g[2]:.4f}")
    assert g[0] < 0.09 and g[1] < 0.26 and g[2] < 0.11
