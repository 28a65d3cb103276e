"""Restart from an HDF5 snapshot (rich_amd.snapshot): a GPU run is written
mid-way, read back and continued the way RICH restarts (extensives formed
anew from the cells, the mesh rebuilt from the stored points).  The
continued GPU run equals the CPU oracle continued from the same file, bit
for bit, for a static periodic mesh and for a moving mesh."""
import numpy as np
import pytest

import cases
import pyoracle as po
from rich_amd import mesh as rmesh
from rich_amd import snapshot

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not snapshot.available(), reason="built without HDF5")]


def test_static_restart(tmp_path):
    from rich_amd.gpu import RichGPU
    m = dict(cases.small_meshes())["periodic"]
    cells = cases.rough_cells(m, seed=12)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    g.init_extensives()
    dts = g.advance(4, integrator=2)
    p = str(tmp_path / "kh.h5")
    snapshot.write(p, m, g.get_cells(), time=float(dts.sum()), cycle=4)
    g.close()
    s = snapshot.read(p)
    assert s["cycle"] == 4
    np.testing.assert_array_equal(s["x"], m.pt_x[:m.n_cells])
    g2 = RichGPU(m, order=2)
    g2.set_cells(s["cells"])
    g2.init_extensives()
    d2 = g2.advance(5, integrator=2)
    o = po.Oracle(m, order=2)
    co, eo, do = o.run(s["cells"], o.init_extensives(s["cells"]), 5, integrator=2)
    np.testing.assert_array_equal(d2, do)
    np.testing.assert_array_equal(g2.get_cells(), co)
    g2.close()


def test_moving_restart(tmp_path):
    from rich_amd.moving import DeviceMovingMesh
    from oracle_drivers import oracle_moving_mesh
    m = cases.sedov_mesh(24)
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    run = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], cases.sedov_cells(m), box)
    for _ in range(4):
        run.step()
    p = str(tmp_path / "sedov.h5")
    snapshot.write(p, run.mesh, run.cells(), time=run.time, cycle=4)
    run.close()
    s = snapshot.read(p)
    a = DeviceMovingMesh(s["x"], s["y"], s["cells"], box)
    b = oracle_moving_mesh(s["x"], s["y"], s["cells"], box)
    for step in range(4):
        assert a.step() == b.step(), step
    np.testing.assert_array_equal(a.x, b.x)
    np.testing.assert_array_equal(a.cells(), b.cells)
    a.close()
