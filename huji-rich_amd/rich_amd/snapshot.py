"""HDF5 snapshots of a 2D run in RICH's layout (lib/librichio.so,
include/richio.h): write_snapshot_to_hdf5 / read_hdf5_snapshot
(hdf5_diagnostics.cpp:237-453), so RICH's restart path, graders and plotting
scripts read these files and this package reads RICH's.

Cells are (4, N) arrays of density, pressure, x and y velocity, as
elsewhere in this package."""
import ctypes
import os

import numpy as np

from . import _ffi

RIO_E = {-1: "bad argument", -2: "HDF5 error", -3: "not a RICH 2D snapshot",
         -4: "size mismatch"}


class _MeshS(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int32)] + [
        (k, ctypes.c_void_p) for k in ("pt_x", "pt_y", "cm_x", "cm_y", "tess_volume", "volume",
                                      "csr_off", "csr_face", "v1x", "v1y", "v2x", "v2y")]


def available():
    """Whether the snapshot library was built and its HDF5 C library loads."""
    if not os.path.exists(os.path.join(_ffi.LIB_DIR, "librichio.so")):
        return False
    try:
        _lib()
    except OSError:
        return False
    return True


def _lib():
    lib = _ffi.load("librichio")
    if not getattr(lib, "_typed", False):
        vp, P = ctypes.c_void_p, ctypes.POINTER
        lib.rio_write_snapshot.argtypes = [ctypes.c_char_p, P(_MeshS), vp, vp, vp, vp,
                                           ctypes.c_double, ctypes.c_int32]
        lib.rio_snapshot_cells.argtypes = [ctypes.c_char_p, P(ctypes.c_int32)]
        lib.rio_read_snapshot.argtypes = [ctypes.c_char_p, ctypes.c_int32] + [vp] * 6 + [
            P(ctypes.c_double), P(ctypes.c_int32)]
        lib.rio_cell_vertices.argtypes = [P(_MeshS), P(ctypes.c_int64), vp, vp, vp,
                                          ctypes.c_int64]
        lib.rio_describe.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]
        lib._typed = True
    return lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"richio: {RIO_E.get(rc, rc)}")


def _mesh(mesh):
    N = mesh.n_cells
    keep = {
        "pt_x": np.ascontiguousarray(mesh.pt_x[:N]), "pt_y": np.ascontiguousarray(mesh.pt_y[:N]),
        "cm_x": np.ascontiguousarray(mesh.cm_x[:N]), "cm_y": np.ascontiguousarray(mesh.cm_y[:N]),
        "tess_volume": np.ascontiguousarray(mesh.tess_volume, dtype=np.float64),
        "volume": np.ascontiguousarray(mesh.volume, dtype=np.float64),
        "csr_off": np.ascontiguousarray(mesh.csr_off, dtype=np.int32),
        "csr_face": np.ascontiguousarray(mesh.csr_face, dtype=np.int32),
        "v1x": np.ascontiguousarray(mesh.v1x), "v1y": np.ascontiguousarray(mesh.v1y),
        "v2x": np.ascontiguousarray(mesh.v2x), "v2y": np.ascontiguousarray(mesh.v2y)}
    s = _MeshS(N, *[keep[k].ctypes.data for k, _ in _MeshS._fields_[1:]])
    return s, keep


def write(path, mesh, cells, time=0.0, cycle=0):
    c = np.ascontiguousarray(cells, dtype=np.float64)
    if c.shape != (4, mesh.n_cells):
        raise ValueError(f"cells must be (4, {mesh.n_cells})")
    s, keep = _mesh(mesh)
    _check(_lib().rio_write_snapshot(os.fsencode(path), ctypes.byref(s),
                                     *[c[k].ctypes.data for k in range(4)],
                                     float(time), int(cycle)))


def read(path):
    """{'x', 'y' (N,), 'cells' (4, N), 'time', 'cycle'} of a snapshot."""
    lib = _lib()
    n = ctypes.c_int32()
    _check(lib.rio_snapshot_cells(os.fsencode(path), ctypes.byref(n)))
    N = n.value
    x, y, c = np.empty(N), np.empty(N), np.empty((4, N))
    t, cy = ctypes.c_double(), ctypes.c_int32()
    _check(lib.rio_read_snapshot(os.fsencode(path), N, x.ctypes.data, y.ctypes.data,
                                 *[c[k].ctypes.data for k in range(4)],
                                 ctypes.byref(t), ctypes.byref(cy)))
    return {"x": x, "y": y, "cells": c, "time": t.value, "cycle": cy.value}


def cell_vertices(mesh):
    """(n_vertices (N,), x_vertices, y_vertices) as a snapshot stores them."""
    lib = _lib()
    s, keep = _mesh(mesh)
    total = ctypes.c_int64()
    _check(lib.rio_cell_vertices(ctypes.byref(s), ctypes.byref(total), None, None, None, 0))
    n, xv, yv = np.empty(mesh.n_cells), np.empty(total.value), np.empty(total.value)
    _check(lib.rio_cell_vertices(ctypes.byref(s), ctypes.byref(total), n.ctypes.data,
                                 xv.ctypes.data, yv.ctypes.data, total.value))
    return n, xv, yv


def describe(path):
    """[(dataset path, class 'f'/'i', bytes, length, chunk, deflate level)]."""
    buf = ctypes.create_string_buffer(1 << 16)
    _check(_lib().rio_describe(os.fsencode(path), buf, len(buf)))
    out = []
    for line in buf.value.decode().splitlines():
        name, cls, size, n, chunk, defl = line.split()
        out.append((name, cls, int(size), int(n), int(chunk), int(defl)))
    return out
