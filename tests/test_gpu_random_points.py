"""An irregular mesh: the Voronoi tessellation of uniformly random points in
a rigid box.  Its cells have 3 to 12 or more faces, so the slope loops run
three chunks of incidences (the lattice meshes of the other tests have at
most 8 faces per cell), and ghost and real neighbours mix within a chunk.
Checked against the oracle bit for bit (plugins and 12 Heun steps), the
mixed-precision step against FP64, and the adjoint by the dot-product
identity against the forward-mode oracle."""
import numpy as np
import pytest

import cases
import pyoracle as po
from parity import assert_parity
from rich_amd import mesh as rmesh
from test_gpu_adjoint import _check as adjoint_check
from test_gpu_mixed import _check as mixed_check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mesh():
    rng = np.random.default_rng(2024)
    n = 3000
    m = rmesh.points(rng.uniform(0.0, 1.0, n), rng.uniform(0.0, 1.0, n), (0.0, 1.0, 0.0, 1.0))
    deg = np.diff(m.csr_off[:m.n_cells + 1])
    print(f"{m.n_cells} cells, faces per cell {deg.min()}..{deg.max()} (mean {deg.mean():.2f})")
    assert deg.max() >= 9 and deg.min() <= 4
    return m


def _cells(m, seed):
    c = cases.rough_cells(m, seed=seed)
    c[0] = np.abs(c[0]) + 0.5
    c[1] = np.abs(c[1]) + 0.5
    return c


def test_random_points_reconstruct_and_flux(mesh):
    from rich_amd.gpu import RichGPU
    cells = _cells(mesh, 1)
    g = RichGPU(mesh, order=2)
    g.set_cells(cells)
    s, L, R = g.reconstruct()
    so, _, Lo, Ro = po.Oracle(mesh, order=2).lgi(cells)
    assert_parity(s, so, what="slopes")
    assert_parity(L, Lo, what="left")
    assert_parity(R, Ro, what="right")
    w = 0.05 * np.random.default_rng(3).standard_normal((2, mesh.n_faces))
    o = po.Oracle(mesh, order=2)
    for face_velocity in (None, w):
        g.set_face_velocity(face_velocity)
        assert_parity(g.flux(), o.flux(cells, face_velocity), what="flux")
    g.close()


@pytest.mark.parametrize("integrator,order", [(2, 2), (1, 1)])
def test_random_points_advance(mesh, integrator, order):
    from rich_amd.gpu import RichGPU
    cells = _cells(mesh, 7)
    o = po.Oracle(mesh, order=order)
    ext = o.init_extensives(cells)
    co, eo, dto = o.run(cells, ext, 12, integrator=integrator)
    g = RichGPU(mesh, order=order)
    g.set_cells(cells)
    g.set_extensives(ext)
    dts = g.advance(12, integrator=integrator)
    assert_parity(dts, dto, what="dt")
    assert_parity(g.get_extensives(), eo, what="extensives")
    _, same = assert_parity(g.get_cells(), co, what="cells")
    assert same == 1.0
    g.close()


def test_random_points_mixed(mesh):
    mixed_check("random_points", mesh, _cells(mesh, 9), 40)


def test_random_points_adjoint(mesh):
    adjoint_check(mesh, _cells(mesh, 11), nsteps=5, every=2)
