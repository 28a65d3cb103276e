"""Moving mesh (rich_amd.moving: hdsim::TimeAdvance2Heun with Lagrangian
points, RoundCells and a Voronoi rebuild every step): the GPU plugins and
the CPU oracle, driven by the same host code on the same meshes, give the
same points, time steps and cells at every step, bit for bit."""
import numpy as np
import pytest

import cases
import pyoracle as po
from oracle_drivers import oracle_moving_mesh

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
