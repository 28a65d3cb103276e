"""Host drivers that run a moving-mesh Heun step entirely on the oracle:
the oracle's plugins (OracleBackend) and the oracle's point motion
(OracleMotion: Lagrangian + RoundCells, CorrectPointsOverShoot,
StationaryBox face velocities), around the same host mesh rebuild the GPU
runs use.  Test infrastructure."""
import numpy as np

import pyoracle as po


class OracleBackend:
    """The plugins of the step on the oracle, one oracle per mesh."""

    def __init__(self, order=2, gamma=5.0 / 3.0):
        self.order, self.gamma = order, gamma

    def set_mesh(self, mesh):
        self.o = po.Oracle(mesh, order=self.order, gamma=self.gamma)

    def init_extensives(self, cells):
        return self.o.init_extensives(cells)

    def cfl(self, cells, fw):
        return self.o.cfl(cells, np.array(fw))

    def flux(self, cells, fw):
        return self.o.flux(cells, np.array(fw))

    def update_extensives(self, flux, dt, ext):
        return self.o.update_extensives(flux, dt, ext)

    def update_cells(self, ext):
        return self.o.update_cells(ext)


class OracleMotion:
    """The point motion of the moving-mesh run on the oracle."""

    def __init__(self, gamma=5.0 / 3.0, chi=0.15, eta=0.02):
        self.gamma, self.chi, self.eta = gamma, chi, eta

    def point_velocity(self, mesh, cells):
        w = po.Oracle(mesh, gamma=self.gamma).point_velocity(cells, self.chi, self.eta)
        return w[0], w[1]

    def point_fix(self, mesh, vx, vy, dt, box):
        w = po.Oracle(mesh, gamma=self.gamma).point_fix(np.vstack([vx, vy]), dt, box)
        return w[0], w[1]

    def face_velocity(self, mesh, vx, vy):
        fw = po.Oracle(mesh, gamma=self.gamma).face_velocity(np.vstack([vx, vy]))
        return fw[0], fw[1]


class OracleLagrangianMotion:
    """Lagrangian point motion (each point with its cell's velocity, no
    fix) with the oracle's StationaryBox face velocities."""

    def point_velocity(self, mesh, cells):
        return np.array(cells[2], dtype=np.float64), np.array(cells[3], dtype=np.float64)

    def point_fix(self, mesh, vx, vy, dt, box):
        return vx, vy

    def face_velocity(self, mesh, vx, vy):
        fw = po.Oracle(mesh).face_velocity(np.vstack([vx, vy]))
        return fw[0], fw[1]


def oracle_moving_mesh(x, y, cells, box, **kw):
    """A MovingMesh whose plugins and point motion are the oracle's."""
    from rich_amd.moving import MovingMesh
    return MovingMesh(x, y, cells, box, OracleBackend(), motion=OracleMotion(), **kw)
