"""The tiles of the fused face-and-update stages (rg_tiles): built at the
first second-order Heun loop of a mesh, with one entry per face and tile that
holds one of its real sides; the fused loop equals the oracle bit for bit on
meshes of several tiles, rigid and periodic, and over odd and even step
counts (the loop alternates two state buffers)."""
import numpy as np
import pytest

import cases
import pyoracle as po
from parity import assert_parity
from rich_amd import mesh as rmesh

pytestmark = pytest.mark.gpu

TILE = 256


def _entries(m):
    """Faces the fused stage solves, counted from the mesh: each face once per
    tile of one of its real sides; and the largest tile's count."""
    N = m.n_cells
    a, b = np.asarray(m.face_first), np.asarray(m.face_second)
    per_tile = np.zeros((N + TILE - 1) // TILE, dtype=np.int64)
    for side in (a, b):
        real = side < N
        np.add.at(per_tile, side[real] // TILE, 1)
    both = (a < N) & (b < N) & (a // TILE == b // TILE)
    np.add.at(per_tile, a[both] // TILE, -1)
    return int(per_tile.sum()), int(per_tile.max())


@pytest.mark.parametrize("bc", [rmesh.RIGID, rmesh.PERIODIC])
@pytest.mark.parametrize("nsteps", [3, 4])
def test_fused_loop_equals_the_oracle(bc, nsteps):
    from rich_amd.gpu import RichGPU
    m = rmesh.lattice(40, 24, (0.0, 1.0, 0.0, 0.6), jitter=2e-2, seed=5, bc=bc)
    cells = cases.rough_cells(m, seed=3)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    g.init_extensives()
    assert g.tiles() == (False, 0, 0)
    dts = g.advance(nsteps, integrator=2)
    fused, n, mx = g.tiles()
    assert fused and (n, mx) == _entries(m) and m.n_cells > 3 * TILE
    o = po.Oracle(m, order=2)
    co, eo, dto = o.run(cells, o.init_extensives(cells), nsteps, integrator=2)
    assert_parity(dts, dto, what="dt")
    np.testing.assert_array_equal(g.get_cells(), co)
    np.testing.assert_array_equal(g.get_extensives(), eo)
    g.close()


def test_mixed_context_builds_its_own_tiles():
    from rich_amd.gpu import RG_MIXED, RichGPU
    m = rmesh.lattice(40, 24, (0.0, 1.0, 0.0, 0.6), jitter=2e-2, seed=5, bc=rmesh.PERIODIC)
    cells = cases.rough_cells(m, seed=3)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    g = RichGPU(m, order=2, precision=RG_MIXED)
    g.set_cells(cells)
    g.init_extensives()
    g.advance(2, integrator=2)
    fused, n, mx = g.tiles()
    # a periodic boundary face is read through its lower-numbered copy, so
    # the entries may differ from the FP64 count, never by more than the
    # boundary faces
    assert fused and abs(n - _entries(m)[0]) <= 2 * m.n_ghosts and 0 < mx <= 2040
    g.close()


def _rough(m, seed=3):
    cells = cases.rough_cells(m, seed=seed)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    return cells


@pytest.mark.parametrize("mixed", [False, True])
def test_fused_and_three_kernel_stages_agree(mixed):
    """rg_use_tiles(0) runs the three-kernel stages: the same bits as the
    fused ones, in FP64 and in mixed precision, also when the switch is
    flipped between two calls (the fused loop left the state in its second
    buffers)."""
    from rich_amd.gpu import RG_FP64, RG_MIXED, RichGPU
    m = rmesh.lattice(40, 24, (0.0, 1.0, 0.0, 0.6), jitter=2e-2, seed=5, bc=rmesh.PERIODIC)
    cells = _rough(m)
    runs = []
    for plan in ((True, True), (False, False), (True, False)):
        g = RichGPU(m, order=2, precision=RG_MIXED if mixed else RG_FP64)
        g.set_cells(cells)
        g.init_extensives()
        dts = []
        for on in plan:
            g.use_tiles(on)
            dts.append(g.advance(3, integrator=2))
            assert g.tiles()[0] == on
        runs.append((np.concatenate(dts), g.get_cells(), g.get_extensives()))
        g.close()
    for got in runs[1:]:
        for a, b in zip(got, runs[0]):
            np.testing.assert_array_equal(a, b)


def test_ensemble_fused_and_three_kernel_stages_agree():
    from rich_amd import ensemble
    from rich_amd.gpu import RichGPU
    ms, cells, _ = ensemble.sedov_ensemble(K=5, n_min=8, n_max=20, seed=3)
    u = ensemble.union(ms)
    runs = []
    for on in (True, False):
        g = RichGPU(u.mesh, order=2)
        g.use_tiles(on)
        g.set_cells(np.concatenate(cells, axis=1))
        g.init_extensives()
        g.set_instances(u.cell_off)
        dts = g.advance_ensemble(8)
        assert g.tiles()[0] == on and u.mesh.n_cells > 2 * TILE
        runs.append((dts, g.get_cells(), g.get_extensives()))
        g.close()
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


def test_adjoint_fused_and_three_kernel_stages_agree():
    """The checkpointing forward and the backward's recompute give the same
    gradient bit for bit with either stage form."""
    from rich_amd.gpu import RichGPU
    m = cases.sedov_mesh(24)
    cells = cases.sedov_cells(m, radius=0.15)
    w = np.random.default_rng(7).standard_normal((4, m.n_cells))
    runs = []
    for on in (True, False):
        g = RichGPU(m, order=2)
        g.use_tiles(on)
        g.set_cells(cells)
        g.init_extensives()
        dts = g.adjoint_forward(12, every=4)
        assert g.tiles()[0] == on and m.n_cells > 2 * TILE
        runs.append((dts, g.get_cells(), g.adjoint_backward(w)))
        g.close()
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


def test_fused_loop_with_face_velocities_equals_the_oracle():
    """The tile kernel reads face velocities through each entry's face id:
    rg_advance with a face-velocity field equals the oracle's run with the same
    field, and the three-kernel stages, bit for bit."""
    from rich_amd.gpu import RichGPU
    m = rmesh.lattice(40, 24, (0.0, 1.0, 0.0, 0.6), jitter=2e-2, seed=8, bc=rmesh.PERIODIC)
    cells = _rough(m, seed=6)
    w = 0.02 * np.random.default_rng(9).standard_normal((2, m.n_faces))
    o = po.Oracle(m, order=2)
    co, eo, dto = o.run(cells, o.init_extensives(cells), 3, integrator=2, w=w)
    runs = []
    for on in (True, False):
        g = RichGPU(m, order=2)
        g.use_tiles(on)
        g.set_cells(cells)
        g.init_extensives()
        g.set_face_velocity(w)
        dts = g.advance(3, integrator=2)
        assert g.tiles()[0] == on
        runs.append((dts, g.get_cells(), g.get_extensives()))
        g.close()
    assert_parity(runs[0][0], dto, what="dt")
    np.testing.assert_array_equal(runs[0][1], co)
    np.testing.assert_array_equal(runs[0][2], eo)
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


def test_shuffled_cell_order_keeps_the_three_kernel_stage():
    """Tiles of cells with little spatial locality read many faces twice; the
    fused stage is then slower than the three-kernel one, and a mesh whose
    tiles would hold more than 1.3 entries per face keeps the latter.  The
    same points in lattice order and shuffled: the first runs fused, the
    second not, and both equal the oracle bit for bit."""
    from rich_amd.gpu import RichGPU
    box = (0.0, 1.0, 0.0, 0.75)
    m0 = rmesh.lattice(48, 36, box, jitter=2e-2, seed=7)
    N = m0.n_cells
    order = np.random.default_rng(4).permutation(N)
    for ids, fused in ((np.arange(N), True), (order, False)):
        m = rmesh.points(np.asarray(m0.pt_x[:N])[ids], np.asarray(m0.pt_y[:N])[ids], box)
        n, _ = _entries(m)
        assert (10 * n <= 13 * m.n_faces) == fused
        cells = _rough(m, seed=6)
        g = RichGPU(m, order=2)
        g.set_cells(cells)
        g.init_extensives()
        dts = g.advance(3, integrator=2)
        assert g.tiles()[0] == fused
        o = po.Oracle(m, order=2)
        co, eo, dto = o.run(cells, o.init_extensives(cells), 3, integrator=2)
        assert_parity(dts, dto, what="dt")
        np.testing.assert_array_equal(g.get_cells(), co)
        np.testing.assert_array_equal(g.get_extensives(), eo)
        g.close()
