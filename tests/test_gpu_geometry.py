"""Cell volumes and centres of mass formed on the device from the uploaded
vertices (rg_mesh geometry NULL): equal, bit for bit, to the host mesh
builder's (which follows CalcCellCM, SlabSymmetry::calcVolume, GetVolume
and the ghost CM rules of VoronoiMesh), and the step on such a context
equals the oracle's."""
import numpy as np
import pytest

import cases
import pyoracle as po
from parity import assert_parity
from rich_amd import mesh as rmesh

pytestmark = pytest.mark.gpu


BOXES = {"rigid": (0.0, 1.2, 0.0, 1.0), "rigid_flip": (-1.0, 1.0, -0.8, 0.8)}


def _meshes():
    out = [(n, m, BOXES.get(n)) for n, m in cases.small_meshes()]
    rng = np.random.default_rng(5)
    x, y = rng.uniform(-1, 1, 2000), rng.uniform(-1, 1, 2000)
    box = (-1.0, 1.0, -1.0, 1.0)
    out.append(("random_points", rmesh.points(x, y, box), box))
    return out


@pytest.mark.parametrize("name,m,box", _meshes(), ids=[n for n, _, _ in _meshes()])
def test_device_geometry_equals_host(name, m, box):
    from rich_amd.gpu import RichGPU
    g = RichGPU(m, order=2, device_geometry=True, box=box)
    v, tv, cx, cy = g.geometry()
    np.testing.assert_array_equal(v, m.volume)
    np.testing.assert_array_equal(tv, m.tess_volume)
    np.testing.assert_array_equal(cx, m.cm_x)
    np.testing.assert_array_equal(cy, m.cm_y)
    cells = cases.rough_cells(m, seed=2)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    o = po.Oracle(m, order=2)
    ext = o.init_extensives(cells)
    co, eo, do = o.run(cells, ext, 5)
    g.set_cells(cells)
    g.init_extensives()
    assert_parity(g.advance(5), do, what="dt")
    assert_parity(g.get_cells(), co, what="cells")
    g.close()


def test_rigid_device_geometry_needs_the_box():
    from rich_amd.gpu import RichGPU, RichGPUError
    m = dict(cases.small_meshes())["rigid"]
    with pytest.raises(RichGPUError):
        RichGPU(m, order=2, device_geometry=True)
