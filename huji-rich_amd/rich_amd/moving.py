"""Second-order steps on a moving mesh in a rigid-wall box:
hdsim::TimeAdvance2Heun (hdsim2d.cpp:323-449) with Lagrangian point motion
and RoundCells (round_cells.cpp:63-134), StationaryBox face velocities and
a host Voronoi rebuild every step.

`MovingMesh` runs the point motion and the mesh rebuild on the host, as in
RICH, and the plugins (CFL, reconstruction and fluxes, extensive and cell
updates) on a backend.  `GpuBackend` calls the HIP library through its C
ABI; any object with the same five methods can stand in for it (the tests
pass the CPU oracle to check the GPU run step by step).

`DeviceMovingMesh` keeps cells, extensives and point velocities on the GPU:
point and face velocities, the overshoot fix, both CFL passes and both
stages run there (rg_moving_begin / rg_moving_end); the host moves the
points and rebuilds the mesh (rg_replace_mesh).  Both give the same values.

Both rebuild the mesh with a `BoxMesher`, which repairs the previous
step's Delaunay triangulation by edge flips where it can.

Every `resort_every` cycles (25, as hdsim: MoveMeshPoints with reorder on
cycle % 25 == 0, hdsim2d.cpp:389-399) the moved points are put in Hilbert
order and the cells, extensives, mid extensives and point velocities follow
them (`hilbert_order`: the textbook curve on a 2^16 grid over the box, not
the reference's adaptive HilbertOrder, so the order can differ from
hdsim's)."""
import time

import numpy as np

from . import mesh as rmesh
from .gpu import RichGPU


class GpuBackend:
    """The plugins on the GPU, one context per mesh."""

    def __init__(self, gamma=5.0 / 3.0, cfl=0.3, order=2):
        self.gamma, self.cfl_number, self.order = gamma, cfl, order
        self.g = None

    def set_mesh(self, mesh):
        if self.g is not None:
            self.g.close()
        self.g = RichGPU(mesh, gamma=self.gamma, cfl=self.cfl_number, order=self.order)

    def init_extensives(self, cells):
        self.g.set_cells(cells)
        self.g.init_extensives()
        return self.g.get_extensives()

    def cfl(self, cells, fw):
        self.g.set_cells(cells)
        self.g.set_face_velocity(fw)
        return self.g.time_step()

    def flux(self, cells, fw):
        self.g.set_cells(cells)
        self.g.set_face_velocity(fw)
        return self.g.flux()

    def update_extensives(self, flux, dt, ext):
        self.g.set_extensives(ext)
        self.g.set_fluxes(flux)
        self.g.update_extensives(dt)
        return self.g.get_extensives()

    def update_cells(self, ext):
        self.g.set_extensives(ext)
        self.g.update_cells()
        return self.g.get_cells()

    def close(self):
        if self.g is not None:
            self.g.close()
            self.g = None


def round_cells_dw(mesh, cells, gamma, chi=0.15, eta=0.02):
    """RoundCells::calc_dw without the cold-flow switch
    (round_cells.cpp:63-76): chi max(c, |v|) (CM - r) / R for cells whose
    point is at least 0.9 eta R from its CM, R = sqrt(GetVolume / pi)."""
    N = mesh.n_cells
    rx, ry = mesh.pt_x[:N], mesh.pt_y[:N]
    sx, sy = mesh.cm_x[:N] - rx, mesh.cm_y[:N] - ry
    d = np.sqrt(sx * sx + sy * sy)
    R = np.sqrt(mesh.tess_volume / np.pi)
    c = np.sqrt(gamma * cells[1] / cells[0])
    vabs = np.sqrt(cells[2] * cells[2] + cells[3] * cells[3])
    c = np.where(c < vabs, vabs, c)  # std::max(c, |v|)
    k = chi * c
    on = ~(d < 0.9 * eta * R)
    return np.where(on, k * sx / R, 0.0), np.where(on, k * sy / R, 0.0)


def correct_overshoot(mesh, vx, vy, dt, box):
    """CorrectPointsOverShoot of a rigid box (round_cells.cpp:14-61): a point
    that would cross a wall within two steps is slowed."""
    x0, x1, y0, y1 = box
    N = mesh.n_cells
    px, py = mesh.pt_x[:N], mesh.pt_y[:N]
    R = np.sqrt(mesh.tess_volume / np.pi)
    inv_dt = 1.0 / dt
    vx, vy = vx.copy(), vy.copy()

    def fix(mask, dist):
        if not mask.any():
            return
        a = np.sqrt(vx[mask] * vx[mask] + vy[mask] * vy[mask])
        f = 0.25 * dist[mask] * inv_dt / a
        f = np.where(R[mask] * 0.1 > dist[mask], f * -0.1, f)
        vx[mask] = vx[mask] * f
        vy[mask] = vy[mask] * f

    fix(vx * dt * 2 + px > x1, x1 - px)
    fix(vx * dt * 2 + px < x0, px - x0)
    fix(vy * dt * 2 + py > y1, y1 - py)
    fix(vy * dt * 2 + py < y0, py - y0)
    return vx, vy


def hilbert_order(x, y, box, bits=16):
    """Indices that sort the points along a Hilbert curve through the box
    (position on the curve of each point's cell in a 2^bits x 2^bits grid;
    ties in the order given)."""
    x0, x1, y0, y1 = box
    n = 1 << bits
    xi = np.clip(((np.asarray(x) - x0) / (x1 - x0) * n).astype(np.int64), 0, n - 1)
    yi = np.clip(((np.asarray(y) - y0) / (y1 - y0) * n).astype(np.int64), 0, n - 1)
    d = np.zeros(len(xi), dtype=np.int64)
    s = n // 2
    while s > 0:
        rx = (xi & s) > 0
        ry = (yi & s) > 0
        d += s * s * ((3 * rx.astype(np.int64)) ^ ry.astype(np.int64))
        # rotate the quadrant
        flip = ~ry & rx
        xi = np.where(flip, n - 1 - xi, xi)
        yi = np.where(flip, n - 1 - yi, yi)
        xi, yi = np.where(~ry, yi, xi), np.where(~ry, xi, yi)
        s //= 2
    return np.argsort(d, kind="stable")


def _resort_order(self):
    """On a re-sort cycle, the Hilbert order of the moved points if it
    differs from the current one (counted in `reorders`), else None."""
    if not self.resort_every or self.cycle % self.resort_every:
        return None
    order = hilbert_order(self.x, self.y, self.box)
    if np.array_equal(order, np.arange(len(order))):
        return None
    self.reorders += 1
    return order


class HostMotion:
    """The point motion of the run on the host: Lagrangian + RoundCells
    point velocities, the overshoot fix and StationaryBox face velocities."""

    def __init__(self, gamma=5.0 / 3.0):
        self.gamma = gamma

    def point_velocity(self, mesh, cells):
        dwx, dwy = round_cells_dw(mesh, cells, self.gamma)
        return cells[2] + dwx, cells[3] + dwy

    def point_fix(self, mesh, vx, vy, dt, box):
        return correct_overshoot(mesh, vx, vy, dt, box)

    def face_velocity(self, mesh, vx, vy):
        return mesh.face_velocity(vx, vy)


class LagrangianMotion:
    """Lagrangian point motion (lagrangian.cpp: each point moves with its
    cell's velocity; PointMotion::ApplyFix leaves the velocities,
    point_motion.cpp:5-10) with StationaryBox face velocities."""

    def point_velocity(self, mesh, cells):
        return np.array(cells[2], dtype=np.float64), np.array(cells[3], dtype=np.float64)

    def point_fix(self, mesh, vx, vy, dt, box):
        return vx, vy

    def face_velocity(self, mesh, vx, vy):
        return mesh.face_velocity(vx, vy)


class MovingMesh:
    """A moving-mesh run: points, cells (4, N) and extensives on the host
    between steps; `step()` advances one TimeAdvance2Heun step (integrator
    2) or one TimeAdvance step (integrator 1).  `motion` supplies the point
    and face velocities (default HostMotion)."""

    def __init__(self, x, y, cells, box, backend, gamma=5.0 / 3.0, motion=None,
                 resort_every=25, integrator=2):
        self.x = np.array(x, dtype=np.float64)
        self.y = np.array(y, dtype=np.float64)
        self.box, self.b, self.gamma = box, backend, gamma
        self.motion = motion if motion is not None else HostMotion(gamma)
        if integrator not in (1, 2):
            raise ValueError("integrator is 1 (TimeAdvance) or 2 (TimeAdvance2Heun)")
        self.integrator = integrator
        self.resort_every, self.cycle, self.reorders = resort_every, 0, 0
        self.mesher = rmesh.BoxMesher(len(self.x), box)
        self.mesh = self.mesher.mesh(self.x, self.y)
        self.b.set_mesh(self.mesh)
        self.cells = np.array(cells, dtype=np.float64)
        self.ext = self.b.init_extensives(self.cells)
        self.time = 0.0
        self.timing = {"mesh_s": 0.0, "motion_s": 0.0, "plugins_s": 0.0}

    _resort_order = _resort_order

    def _plug(self, fn, *a):
        t = time.perf_counter()
        out = fn(*a)
        self.timing["plugins_s"] += time.perf_counter() - t
        return out

    def step(self):
        if self.integrator == 1:
            return self._step_euler()
        b, T = self.b, self.timing
        t = time.perf_counter()
        # Lagrangian + RoundCells point velocities, StationaryBox face velocities
        mo = self.motion
        vx, vy = mo.point_velocity(self.mesh, self.cells)
        fw = mo.face_velocity(self.mesh, vx, vy)
        T["motion_s"] += time.perf_counter() - t
        dt = self._plug(b.cfl, self.cells, fw)
        t = time.perf_counter()
        vx, vy = mo.point_fix(self.mesh, vx, vy, dt, self.box)  # ApplyFix
        fw = mo.face_velocity(self.mesh, vx, vy)
        T["motion_s"] += time.perf_counter() - t
        dt = self._plug(b.cfl, self.cells, fw)
        # stage 1 on the old mesh; ZeroForce adds dt*0 (hydrodynamics_2d.cpp:348-372)
        flux = self._plug(b.flux, self.cells, fw)
        mid = self.ext.copy()
        mid[0] += dt * 0.0
        mid[2] += 0.0 * dt
        mid[3] += 0.0 * dt
        mid[1] += dt * 0.0
        mid = self._plug(b.update_extensives, flux, dt, mid)
        # MoveMeshPoints (hydrodynamics_2d.cpp:218-247), the Hilbert re-sort
        # every resort_every cycles, and the rebuild
        t = time.perf_counter()
        self.x = self.x + dt * vx
        self.y = self.y + dt * vy
        order = self._resort_order()
        if order is not None:
            self.x, self.y, vx, vy = self.x[order], self.y[order], vx[order], vy[order]
            self.cells, self.ext, mid = self.cells[:, order], self.ext[:, order], mid[:, order]
            self.mesher.permute(order)
        self.mesh = self.mesher.mesh(self.x, self.y)
        T["mesh_s"] += time.perf_counter() - t
        self._plug(b.set_mesh, self.mesh)
        mid_cells = self._plug(b.update_cells, mid)
        t = time.perf_counter()
        fw = mo.face_velocity(self.mesh, vx, vy)
        T["motion_s"] += time.perf_counter() - t
        # stage 2 on the new mesh
        flux = self._plug(b.flux, mid_cells, fw)
        mid[0] += dt * 0.0
        mid[2] += 0.0 * dt
        mid[3] += 0.0 * dt
        mid[1] += dt * 0.0
        mid = self._plug(b.update_extensives, flux, dt, mid)
        # average_extensive (hdsim2d.cpp:311-320) and the cell update
        ext = np.empty_like(mid)
        ext[0] = 0.5 * (mid[0] + self.ext[0])
        ext[1] = 0.5 * (mid[1] + self.ext[1])
        ext[2] = (mid[2] + self.ext[2]) * 0.5
        ext[3] = (mid[3] + self.ext[3]) * 0.5
        self.ext = ext
        self.cells = self._plug(b.update_cells, ext)
        self.time += dt
        self.cycle += 1
        return dt


    def _step_euler(self):
        """hdsim::TimeAdvance (hdsim2d.cpp:149-241): point and face
        velocities, dt, ApplyFix, the fluxes and the extensive update on the
        old mesh, ExternalForceContribution with ZeroForce, MoveMeshPoints
        with the re-sort every resort_every cycles, the rebuild and the cell
        update on the new mesh."""
        b, T = self.b, self.timing
        t = time.perf_counter()
        mo = self.motion
        vx, vy = mo.point_velocity(self.mesh, self.cells)
        fw = mo.face_velocity(self.mesh, vx, vy)
        T["motion_s"] += time.perf_counter() - t
        dt = self._plug(b.cfl, self.cells, fw)
        t = time.perf_counter()
        vx, vy = mo.point_fix(self.mesh, vx, vy, dt, self.box)
        fw = mo.face_velocity(self.mesh, vx, vy)
        T["motion_s"] += time.perf_counter() - t
        flux = self._plug(b.flux, self.cells, fw)
        ext = self._plug(b.update_extensives, flux, dt, self.ext)
        ext[0] += dt * 0.0
        ext[2] += 0.0 * dt
        ext[3] += 0.0 * dt
        ext[1] += dt * 0.0
        t = time.perf_counter()
        self.x = self.x + dt * vx
        self.y = self.y + dt * vy
        order = self._resort_order()
        if order is not None:
            self.x, self.y = self.x[order], self.y[order]
            self.cells, ext = self.cells[:, order], ext[:, order]
            self.mesher.permute(order)
        self.mesh = self.mesher.mesh(self.x, self.y)
        T["mesh_s"] += time.perf_counter() - t
        self._plug(b.set_mesh, self.mesh)
        self.ext = ext
        self.cells = self._plug(b.update_cells, ext)
        self.time += dt
        self.cycle += 1
        return dt


class DeviceMovingMesh:
    """A moving-mesh run with the state on the GPU; `step()` advances one
    TimeAdvance2Heun step, `cells()` reads the cells back.  `fix_box` is the
    box of RoundCells' overshoot fix (default: the mesh's box, RoundCells
    built with the outer boundary; RoundCells built without one fixes
    against [-1, 1]^2, round_cells.cpp:10-11)."""

    _resort_order = _resort_order

    def __init__(self, x, y, cells, box, gamma=5.0 / 3.0, cfl=0.3, chi=0.15, eta=0.02,
                 device_geometry=True, resort_every=25, fix_box=None):
        self.x = np.array(x, dtype=np.float64)
        self.y = np.array(y, dtype=np.float64)
        self.box, self.chi, self.eta = tuple(float(v) for v in box), chi, eta
        self.fix_box = self.box if fix_box is None else tuple(float(v) for v in fix_box)
        self.resort_every, self.cycle, self.reorders = resort_every, 0, 0
        self.mesher = rmesh.BoxMesher(len(self.x), box)
        self.mesh = self.mesher.mesh(self.x, self.y)
        # volumes and centres of mass formed on the device from each mesh's
        # vertices (not uploaded)
        self.g = RichGPU(self.mesh, gamma=gamma, cfl=cfl, order=2,
                         device_geometry=device_geometry, box=self.box)
        self.g.set_cells(np.asarray(cells, dtype=np.float64))
        self.g.init_extensives()
        self.time = 0.0
        self.timing = {"mesh_s": 0.0, "begin_s": 0.0, "replace_s": 0.0, "end_s": 0.0}

    def step(self):
        T = self.timing
        t0 = time.perf_counter()
        dt, vx, vy = self.g.moving_begin(self.chi, self.eta, self.fix_box)
        t1 = time.perf_counter()
        # MoveMeshPoints (hydrodynamics_2d.cpp:218-247), the Hilbert re-sort
        # every resort_every cycles (the state follows on the device), and
        # the rebuild
        self.x = self.x + dt * vx
        self.y = self.y + dt * vy
        order = self._resort_order()
        if order is not None:
            self.x, self.y = self.x[order], self.y[order]
            self.g.permute_cells(order)
            self.mesher.permute(order)
        self.mesh = self.mesher.mesh(self.x, self.y)
        t2 = time.perf_counter()
        self.g.replace_mesh(self.mesh)
        t3 = time.perf_counter()
        self.g.moving_end()
        t4 = time.perf_counter()
        T["begin_s"] += t1 - t0
        T["mesh_s"] += t2 - t1
        T["replace_s"] += t3 - t2
        T["end_s"] += t4 - t3
        self.time += dt
        self.cycle += 1
        return dt

    def cells(self):
        return self.g.get_cells()

    def extensives(self):
        return self.g.get_extensives()

    def close(self):
        self.g.close()
