"""Moving mesh (rich_amd.moving: hdsim::TimeAdvance2Heun with Lagrangian
points, RoundCells and a Voronoi rebuild every step): the GPU plugins and
the CPU oracle, driven by the same host code on the same meshes, give the
same points, time steps and cells at every step, bit for bit."""
import numpy as np
import pytest

import cases
import pyoracle as po
from oracle_drivers import OracleBackend, oracle_moving_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [20, 32])
def test_moving_sedov_gpu_equals_oracle(n):
    from rich_amd.moving import GpuBackend, MovingMesh
    m = cases.sedov_mesh(n)
    c = cases.sedov_cells(m)
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    runs = [MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, GpuBackend()),
            oracle_moving_mesh(m.pt_x[:N], m.pt_y[:N], c, box)]
    moved = 0.0
    for step in range(8):
        dts = [r.step() for r in runs]
        assert dts[0] == dts[1], step
        np.testing.assert_array_equal(runs[0].x, runs[1].x)
        np.testing.assert_array_equal(runs[0].y, runs[1].y)
        np.testing.assert_array_equal(runs[0].cells, runs[1].cells)
        np.testing.assert_array_equal(runs[0].ext, runs[1].ext)
    moved = np.max(np.hypot(runs[0].x - m.pt_x[:N], runs[0].y - m.pt_y[:N]))
    print(f"n={n}: points moved up to {moved:.3e}; faces now {runs[0].mesh.n_faces} "
          f"(initially {m.n_faces}); t = {runs[0].time:.4e}")
    assert moved > 0
    runs[0].b.close()


@pytest.mark.parametrize("n", [20, 32])
def test_moving_device_resident_equals_oracle(n):
    """rg_moving_begin / rg_replace_mesh / rg_moving_end (point and face
    velocities, the overshoot fix and both stages on the device) against the
    host-driven run over the oracle (plugins and point motion): same points,
    dt and state every step."""
    from rich_amd.moving import DeviceMovingMesh
    m = cases.sedov_mesh(n)
    c = cases.sedov_cells(m)
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box)
    ref = oracle_moving_mesh(m.pt_x[:N], m.pt_y[:N], c, box)
    for step in range(8):
        dt_dev, dt_ref = dev.step(), ref.step()
        print(f"n={n} step={step} dt={dt_dev:.6e}", flush=True)
        assert dt_dev == dt_ref, step
        np.testing.assert_array_equal(dev.x, ref.x)
        np.testing.assert_array_equal(dev.y, ref.y)
        np.testing.assert_array_equal(dev.cells(), ref.cells)
        np.testing.assert_array_equal(dev.extensives(), ref.ext)
    assert dev.mesh.n_faces != m.n_faces or np.any(dev.x != m.pt_x[:N])
    dev.close()


def test_moving_device_rejects_bad_use():
    """rg_moving_end needs a preceding rg_moving_begin; rg_replace_mesh keeps
    the cell count."""
    from rich_amd.gpu import RichGPU, RichGPUError
    m = cases.sedov_mesh(16)
    g = RichGPU(m, order=2)
    g.set_cells(cases.sedov_cells(m))
    g.init_extensives()
    with pytest.raises(RichGPUError):
        g.moving_end()
    other = cases.sedov_mesh(12)
    with pytest.raises(ValueError):
        g.replace_mesh(other)
    g.close()


def test_permute_cells():
    """rg_permute_cells: new cell i = old cell order[i] for cells and
    extensives; anything but a permutation of [0, N) is refused."""
    from rich_amd.gpu import RichGPU, RichGPUError
    m = cases.sedov_mesh(16)
    c = cases.rough_cells(m, seed=3)
    c[0] = np.abs(c[0]) + 0.5
    c[1] = np.abs(c[1]) + 0.5
    g = RichGPU(m, order=2)
    g.set_cells(c)
    g.init_extensives()
    x = g.get_extensives()
    order = np.random.default_rng(1).permutation(m.n_cells)
    g.permute_cells(order)
    np.testing.assert_array_equal(g.get_cells(), c[:, order])
    np.testing.assert_array_equal(g.get_extensives(), x[:, order])
    bad = order.copy()
    bad[0] = bad[1]
    with pytest.raises(RichGPUError):
        g.permute_cells(bad)
    g.close()


@pytest.mark.parametrize("device", [False, True])
def test_resorted_moving_run_equals_oracle(device):
    """Hilbert re-sort every 3 cycles (hdsim does it every 25): the GPU runs,
    host-driven and device-resident (rg_permute_cells moves cells,
    extensives, mid-step extensives and point velocities on the device),
    equal the oracle-driven run that re-sorts on the host, bit for bit."""
    from rich_amd.moving import DeviceMovingMesh, GpuBackend, MovingMesh
    m, c = cases.smooth_box_case()
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    if device:
        run = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, resort_every=3)
    else:
        run = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, GpuBackend(), resort_every=3)
    ref = oracle_moving_mesh(m.pt_x[:N], m.pt_y[:N], c, box, resort_every=3)
    for step in range(10):
        dt, dt_ref = run.step(), ref.step()
        print(f"device={device} step={step} dt={dt:.6e}", flush=True)
        assert dt == dt_ref, step
        np.testing.assert_array_equal(run.x, ref.x)
        np.testing.assert_array_equal(run.y, ref.y)
        cells = run.cells() if device else run.cells
        ext = run.extensives() if device else run.ext
        np.testing.assert_array_equal(cells, ref.cells)
        np.testing.assert_array_equal(ext, ref.ext)
    assert run.reorders == ref.reorders > 0
    if device:
        run.close()
    else:
        run.b.close()


def test_conservation_lagrangian_gpu_equals_oracle():
    """The reference's conservation_lagrangian run (30 x 30 Cartesian grid,
    Lagrangian points, SimpleFluxCalculator, TimeAdvance, to t = 0.05) with
    the GPU plugins: every step equals the oracle-driven run bit for bit, and
    the totals pass the reference's grader (equal at six significant
    digits) with mass and energy within 1e-13."""
    from rich_amd.moving import GpuBackend, LagrangianMotion
    from oracle_drivers import OracleBackend, OracleLagrangianMotion
    from test_conservation_lagrangian import check_totals, lagrangian_run
    gpu = lagrangian_run(GpuBackend(order=1), LagrangianMotion())
    ref = lagrangian_run(OracleBackend(order=1), OracleLagrangianMotion())
    totals = [gpu.ext.sum(axis=1)]
    step = 0
    while ref.time < 0.05:
        dt, dt_ref = gpu.step(), ref.step()
        assert dt == dt_ref, step
        np.testing.assert_array_equal(gpu.x, ref.x)
        np.testing.assert_array_equal(gpu.y, ref.y)
        np.testing.assert_array_equal(gpu.cells, ref.cells)
        np.testing.assert_array_equal(gpu.ext, ref.ext)
        totals.append(gpu.ext.sum(axis=1))
        step += 1
    assert step > 20 and gpu.reorders == ref.reorders >= 1
    check_totals(totals)
    gpu.b.close()


def test_riemann_messy_grid_gpu_equals_oracle():
    """The reference's riemann_profiles_messy_grid run (Lagrangian +
    RoundCells points, SimpleFluxCalculator, TimeAdvance, to t = 0.225)
    with the GPU plugins: every step equals the oracle-driven run bit for
    bit, and the grader's goodness of fit passes."""
    from rich_amd.moving import GpuBackend, HostMotion
    from test_riemann_messy_grid import (REFERENCE_FIX_BOX, OracleMotionReferenceFix, grade,
                                         messy_grid_run)

    class HostMotionReferenceFix(HostMotion):
        def point_fix(self, mesh, vx, vy, dt, box):
            return super().point_fix(mesh, vx, vy, dt, REFERENCE_FIX_BOX)

    gpu = messy_grid_run(GpuBackend(order=1), HostMotionReferenceFix())
    ref = messy_grid_run(OracleBackend(order=1), OracleMotionReferenceFix())
    step = 0
    while ref.time < 0.225:
        dt, dt_ref = gpu.step(), ref.step()
        assert dt == dt_ref, step
        np.testing.assert_array_equal(gpu.x, ref.x)
        np.testing.assert_array_equal(gpu.y, ref.y)
        np.testing.assert_array_equal(gpu.cells, ref.cells)
        step += 1
    gd, gp, gv = grade(gpu)
    assert gd < 0.1 and gp < 0.11 and gv < 0.11
    gpu.b.close()


def test_sedov_taylor_grid_gpu_equals_oracle():
    """The reference's sedov_taylor_messy_grid run (50 x 50 Cartesian grid,
    Lagrangian points, SimpleFluxCalculator, TimeAdvance, 100 steps) with
    the GPU plugins: every step equals the oracle-driven run bit for bit,
    and the grader's comparison with the Sedov-Taylor profiles passes."""
    from rich_amd.moving import GpuBackend, LagrangianMotion
    from oracle_drivers import OracleLagrangianMotion
    from test_sedov_taylor_grid import grade, sedov_grid_run
    gpu = sedov_grid_run(GpuBackend(order=1), LagrangianMotion())
    ref = sedov_grid_run(OracleBackend(order=1), OracleLagrangianMotion())
    for step in range(100):
        dt, dt_ref = gpu.step(), ref.step()
        assert dt == dt_ref, step
        np.testing.assert_array_equal(gpu.x, ref.x)
        np.testing.assert_array_equal(gpu.y, ref.y)
        np.testing.assert_array_equal(gpu.cells, ref.cells)
    front, (ld, lp, lv) = grade(gpu)
    print(f"shock at r = {front:.4f}: L1 {ld:.4f} {lp:.4f} {lv:.4f}")
    assert ld < 0.5 and lp < 14 and lv < 1.5
    gpu.b.close()


def test_simple_waves_device_equals_oracle():
    """The reference's simple_waves_plm_2 run (second order, Lagrangian +
    RoundCells, TimeAdvance2Heun, to t = 0.02) on the device-resident
    driver: every step equals the oracle-driven run bit for bit, and the
    grader's comparison with the simple-wave solution passes."""
    from rich_amd.moving import DeviceMovingMesh, MovingMesh
    from test_simple_waves import (FIX_BOX, GAMMA, OracleMotionDefaultFix, grade,
                                   simple_wave_mesh)
    m, c = simple_wave_mesh()
    N = m.n_cells
    box = (0.0, 1.0, 0.0, 1.0)
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, gamma=GAMMA, fix_box=FIX_BOX)
    ref = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, OracleBackend(order=2, gamma=GAMMA),
                     motion=OracleMotionDefaultFix())
    step = 0
    while ref.time < 0.02:
        dt, dt_ref = dev.step(), ref.step()
        assert dt == dt_ref, step
        np.testing.assert_array_equal(dev.x, ref.x)
        np.testing.assert_array_equal(dev.y, ref.y)
        np.testing.assert_array_equal(dev.cells(), ref.cells)
        step += 1
    g = grade(m.pt_x[:N], c, dev.x, dev.cells(), dev.time)
    assert g[0] < 0.09 and g[1] < 0.26 and g[2] < 0.11
    dev.close()


def test_riemann_messy_grid_second_order_device_equals_oracle():
    """The reference's riemann_profiles_messy_grid_2 run (second order,
    Lagrangian + RoundCells, TimeAdvance2Heun, to t = 0.225) on the
    device-resident driver: equal to the oracle-driven run bit for bit at
    every step, and the grader's bounds hold."""
    from rich_amd import mesh as rmesh
    from rich_amd.moving import DeviceMovingMesh, MovingMesh
    from test_riemann_messy_grid import REFERENCE_FIX_BOX, OracleMotionReferenceFix, grade_state
    m = rmesh.lattice(30, 30, (0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=1, bc=rmesh.RIGID)
    N = m.n_cells
    x = m.pt_x[:N]
    c = np.vstack([np.ones(N), np.where(x < 0.5, 2.0, 1.0), np.zeros(N), np.zeros(N)])
    box = (0.0, 1.0, 0.0, 1.0)
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, fix_box=REFERENCE_FIX_BOX)
    ref = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, OracleBackend(order=2),
                     motion=OracleMotionReferenceFix(), integrator=2)
    step = 0
    while ref.time < 0.225:
        dt, dt_ref = dev.step(), ref.step()
        assert dt == dt_ref, step
        np.testing.assert_array_equal(dev.x, ref.x)
        np.testing.assert_array_equal(dev.cells(), ref.cells)
        step += 1
    gd, gp, gv = grade_state(dev.x, dev.time, dev.cells())
    assert gd < 0.067 and gp < 0.053 and gv < 0.062
    dev.close()


def test_gresho_vortex_device_equals_oracle():
    """The reference's gresho_vortex_2o run on the device-resident driver:
    every step equals the oracle-driven run bit for bit, and the grader's L1
    bounds hold."""
    from rich_amd.moving import DeviceMovingMesh, MovingMesh
    from test_gresho_moving import (FIX_BOX, KNOWN, OracleMotionDefaultFix, check,
                                    gresho_mesh, l1_errors)
    m, c = gresho_mesh()
    N = m.n_cells
    box = tuple(KNOWN["box"])
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, fix_box=FIX_BOX)
    ref = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, OracleBackend(order=2),
                     motion=OracleMotionDefaultFix(), integrator=2)
    step = 0
    while ref.time < KNOWN["t_end"]:
        dt, dt_ref = dev.step(), ref.step()
        assert dt == dt_ref, step
        np.testing.assert_array_equal(dev.x, ref.x)
        np.testing.assert_array_equal(dev.cells(), ref.cells)
        step += 1
    check(l1_errors(dev.x, dev.y, dev.cells()))
    dev.close()
