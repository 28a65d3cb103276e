"""HDF5 snapshots (rich_amd.snapshot, lib/librichio.so) in RICH's 2D layout
(write_snapshot_to_hdf5 / read_hdf5_snapshot, hdf5_diagnostics.cpp:237-453):
exact round trips, the dataset layout of write_std_vector_to_hdf5
(hdf5_utils.hpp:29-51), and each cell's polygon (ConvexHull,
ConvexHull.cpp:30-58) checked against the mesh's own volumes and centres of
mass by the shoelace formula."""
import os

import numpy as np
import pytest

from rich_amd import mesh as rmesh
from rich_amd import snapshot

pytestmark = pytest.mark.skipif(not snapshot.available(), reason="built without HDF5")

BOX = (-1.0, 1.0, -1.0, 1.0)


def _cells(N, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.5, 2, N), rng.uniform(0.1, 3, N),
                     rng.standard_normal(N), rng.standard_normal(N)])


@pytest.mark.parametrize("bc", [rmesh.RIGID, rmesh.PERIODIC])
def test_round_trip_is_exact(tmp_path, bc):
    m = rmesh.lattice(24, 18, BOX, jitter=1e-2, seed=3, bc=bc)
    c = _cells(m.n_cells, 1)
    p = str(tmp_path / "snap.h5")
    snapshot.write(p, m, c, time=0.125, cycle=42)
    s = snapshot.read(p)
    np.testing.assert_array_equal(s["cells"], c)
    np.testing.assert_array_equal(s["x"], m.pt_x[:m.n_cells])
    np.testing.assert_array_equal(s["y"], m.pt_y[:m.n_cells])
    assert s["time"] == 0.125 and s["cycle"] == 42


def test_layout_matches_the_reference_writer(tmp_path):
    """Names, classes, lengths, chunking capped at 100000 and deflate 6."""
    m = rmesh.lattice(400, 260, BOX, jitter=1e-3, seed=5)
    N = m.n_cells
    p = str(tmp_path / "big.h5")
    snapshot.write(p, m, _cells(N, 2), time=1.5, cycle=7)
    d = {row[0]: row[1:] for row in snapshot.describe(p)}
    nv, xv, _ = snapshot.cell_vertices(m)
    V = len(xv)
    expect = {"/time": ("f", 8, 1), "/cycle": ("i", 4, 1), "/Volumes": ("f", 8, N),
              "/geometry/x_coordinate": ("f", 8, N), "/geometry/y_coordinate": ("f", 8, N),
              "/geometry/x_vertices": ("f", 8, V), "/geometry/y_vertices": ("f", 8, V),
              "/geometry/n_vertices": ("f", 8, N),
              "/hydrodynamic/density": ("f", 8, N), "/hydrodynamic/pressure": ("f", 8, N),
              "/hydrodynamic/x_velocity": ("f", 8, N), "/hydrodynamic/y_velocity": ("f", 8, N)}
    assert set(d) == set(expect)
    for k, (cls, size, n) in expect.items():
        assert d[k][:3] == (cls, size, n), k
        assert d[k][3] == min(n, 100000) and d[k][4] == 6, k
    assert N > 100000 and int(nv.sum()) == V


@pytest.mark.parametrize("shake", [0.0, 0.3])
def test_cell_polygons(shake):
    """Vertices in angle order about the CM, one per distinct corner: the
    shoelace area is the cell volume and the polygon centroid its CM."""
    m0 = rmesh.lattice(20, 20, BOX, jitter=1e-3, seed=8)
    N = m0.n_cells
    rng = np.random.default_rng(4)
    x = np.clip(m0.pt_x[:N] + shake * 0.1 * rng.uniform(-1, 1, N), -0.99, 0.99)
    y = np.clip(m0.pt_y[:N] + shake * 0.1 * rng.uniform(-1, 1, N), -0.99, 0.99)
    m = m0 if shake == 0 else rmesh.points(x, y, BOX)
    nv, xv, yv = snapshot.cell_vertices(m)
    off = np.concatenate([[0], np.cumsum(nv.astype(int))])
    for i in range(N):
        px, py = xv[off[i]:off[i + 1]], yv[off[i]:off[i + 1]]
        ang = np.arctan2(py - m.cm_y[i], px - m.cm_x[i])
        assert np.all(np.diff(ang) > 0)
        cr = px * np.roll(py, -1) - np.roll(px, -1) * py
        a = 0.5 * cr.sum()
        assert abs(a - m.volume[i]) <= 1e-12 * m.volume[i]
        cx = ((px + np.roll(px, -1)) * cr).sum() / (6 * a)
        cy = ((py + np.roll(py, -1)) * cr).sum() / (6 * a)
        assert abs(cx - m.cm_x[i]) < 1e-12 and abs(cy - m.cm_y[i]) < 1e-12
        deg = m.csr_off[i + 1] - m.csr_off[i]
        assert nv[i] <= deg


def test_errors(tmp_path):
    m = rmesh.lattice(8, 8, BOX, jitter=1e-3, seed=1)
    with pytest.raises(ValueError):
        snapshot.write(str(tmp_path / "x.h5"), m, np.zeros((4, 3)))
    with pytest.raises(RuntimeError):
        snapshot.read(str(tmp_path / "missing.h5"))
    p = str(tmp_path / "s.h5")
    snapshot.write(p, m, _cells(m.n_cells, 3))
    import ctypes
    lib = snapshot._lib()
    buf = np.empty(m.n_cells + 1)
    rc = lib.rio_read_snapshot(os.fsencode(p), m.n_cells + 1, buf.ctypes.data, None, None, None,
                               None, None, None, None)
    assert rc == -4
    assert lib.rio_snapshot_cells(os.fsencode(p), ctypes.byref(ctypes.c_int32())) == 0
