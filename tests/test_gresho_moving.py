"""The reference's gresho_vortex_2o run as it is set up (tests/newtonian/
two_dimensional/gresho_vortex_2o/test.cpp:140-231, test.py:115-118): a
30 x 30 Cartesian grid in the rigid box [-0.5, 0.5]^2, Lagrangian +
RoundCells points (RoundCells without an outer boundary: overshoot fix
against [-1, 1]^2), LinearGaussImproved + ModularFluxCalculator,
TimeAdvance2Heun, the Gresho vortex at each cell's CM, run while
t < 0.003.  The grader's L1 errors at the final points (density 1, the
vortex's pressure and azimuthal velocity, no radial velocity) must stay
below 2e-4, 2e-3, 5e-4 and 1e-3.  (test_oracle_scheme checks the same
bounds on a static mesh.)"""
import json
import os

import numpy as np

from oracle_drivers import OracleBackend, OracleMotion
from rich_amd import mesh as rmesh
from rich_amd.moving import MovingMesh

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))["gresho_vortex_2o"]
FIX_BOX = (-1.0, 1.0, -1.0, 1.0)


def vq2r(r):
    return np.where(r < 0.2, 5.0, np.where(r > 0.4, 0.0, 2.0 / np.maximum(r, 1e-300) - 5.0))


def pressure(r):
    return np.where(r < 0.2, 5 + 12.5 * r ** 2, np.where(
        r > 0.4, 3 + 4 * np.log(2.0),
        9 + 12.5 * r ** 2 - 20 * r + 4 * np.log(np.maximum(r, 1e-300) / 0.2)))


def vphi(r):
    return np.where(r < 0.2, 5 * r, np.where(r > 0.4, 0.0, 2 - 5 * r))


def gresho_mesh():
    n = KNOWN["n"]
    m = rmesh.lattice(n, n, tuple(KNOWN["box"]), jitter=0.0, seed=0, bc=rmesh.RIGID)
    N = m.n_cells
    cx, cy = m.cm_x[:N], m.cm_y[:N]
    r = np.hypot(cx, cy)
    return m, np.vstack([np.ones(N), pressure(r), -vq2r(r) * cy, vq2r(r) * cx])


class OracleMotionDefaultFix(OracleMotion):
    def point_fix(self, mesh, vx, vy, dt, box):
        return super().point_fix(mesh, vx, vy, dt, FIX_BOX)


def l1_errors(x, y, c):
    rr = np.hypot(x, y)
    vf = (c[3] * x - c[2] * y) / rr
    vr = (c[3] * y + c[2] * x) / rr
    return (float(np.mean(np.abs(c[0] - 1))), float(np.mean(np.abs(c[1] - pressure(rr)))),
            float(np.mean(np.abs(vr))), float(np.mean(np.abs(vf - vphi(rr)))))


def check(err):
    assert err[0] < KNOWN["l1_density"] and err[1] < KNOWN["l1_pressure"]
    assert err[2] < KNOWN["l1_radial_velocity"] and err[3] < KNOWN["l1_azimuthal_velocity"]


def test_gresho_vortex_moving_oracle():
    m, c = gresho_mesh()
    N = m.n_cells
    box = tuple(KNOWN["box"])
    run = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, OracleBackend(order=2),
                     motion=OracleMotionDefaultFix(), integrator=2)
    while run.time < KNOWN["t_end"]:
        run.step()
    err = l1_errors(run.x, run.y, run.cells)
    print(f"t = {run.time:.5f}: L1 {err}")
    check(err)
