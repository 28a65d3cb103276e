"""Flat (structure-of-arrays) Voronoi meshes, as a RICH VoronoiMesh exposes
them through its Tessellation accessors, built for jittered lattices by
lib/librichmesh.so (include/richmesh.h)."""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _ffi

RIGID = 0
PERIODIC = 1


class _LatticeParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("x0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y0", ctypes.c_double), ("y1", ctypes.c_double),
                ("jitter", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("bc", ctypes.c_int32), ("flip_seed", ctypes.c_int32)]


class _PointsParams(ctypes.Structure):
    _fields_ = [("n_real", ctypes.c_int32), ("n_ghost", ctypes.c_int32),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("ghost_orig", ctypes.c_void_p), ("ghost_wall", ctypes.c_void_p),
                ("n_tri", ctypes.c_int32), ("tri", ctypes.c_void_p),
                ("x0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y0", ctypes.c_double), ("y1", ctypes.c_double)]


class _BoxPointsParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("x", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("x0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y0", ctypes.c_double), ("y1", ctypes.c_double),
                ("margin", ctypes.c_double)]


def _lib():
    lib = _ffi.load("librichmesh")
    if not getattr(lib, "_typed", False):
        lib.rm_lattice_build.restype = ctypes.c_void_p
        lib.rm_lattice_build.argtypes = [ctypes.POINTER(_LatticeParams)]
        lib.rm_points_build.restype = ctypes.c_void_p
        lib.rm_points_build.argtypes = [ctypes.POINTER(_PointsParams)]
        lib.rm_free.argtypes = [ctypes.c_void_p]
        lib.rm_sizes.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int32)] * 4
        lib.rm_export.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 17
        lib.rm_face_velocity.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        lib.rm_mover_create.restype = ctypes.c_void_p
        lib.rm_mover_create.argtypes = [ctypes.POINTER(_BoxPointsParams)]
        lib.rm_mover_mesh.restype = ctypes.c_void_p
        lib.rm_mover_mesh.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.rm_mover_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_int32)]
        lib.rm_mover_free.argtypes = [ctypes.c_void_p]
        lib.rm_mover_permute.restype = ctypes.c_int
        lib.rm_mover_permute.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.rm_box_points_build.restype = ctypes.c_void_p
        lib.rm_box_points_build.argtypes = [ctypes.POINTER(_BoxPointsParams)]
        lib.rm_delaunay.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        lib._typed = True
    return lib


@dataclass
class Mesh:
    """Flat mesh.  Ghost points are numbered n_cells .. n_cells+n_ghosts-1."""
    n_cells: int
    n_faces: int
    n_ghosts: int
    bc: int
    face_first: np.ndarray
    face_second: np.ndarray
    v1x: np.ndarray
    v1y: np.ndarray
    v2x: np.ndarray
    v2y: np.ndarray
    pt_x: np.ndarray
    pt_y: np.ndarray
    cm_x: np.ndarray
    cm_y: np.ndarray
    volume: np.ndarray
    tess_volume: np.ndarray
    csr_off: np.ndarray
    csr_face: np.ndarray
    ghost_orig: np.ndarray
    lat_i: np.ndarray = None
    lat_j: np.ndarray = None
    _handle: object = None

    def orig(self, ids):
        ids = np.asarray(ids)
        out = ids.copy()
        g = ids >= self.n_cells
        out[g] = self.ghost_orig[ids[g] - self.n_cells]
        return out

    def face_velocity(self, wx, wy):
        """Face velocities for point velocities (StationaryBox /
        PeriodicEdgeVelocities)."""
        if self._handle is None:
            raise RuntimeError("face_velocity needs a mesh built by lattice()")
        fwx = np.empty(self.n_faces)
        fwy = np.empty(self.n_faces)
        wx = np.ascontiguousarray(wx, dtype=np.float64)
        wy = np.ascontiguousarray(wy, dtype=np.float64)
        _lib().rm_face_velocity(self._handle.value, _ffi.ptr(wx), _ffi.ptr(wy),
                                _ffi.ptr(fwx), _ffi.ptr(fwy))
        return fwx, fwy


class _Handle:
    def __init__(self, value):
        self.value = value

    def __del__(self):
        try:
            _lib().rm_free(self.value)
        except Exception:
            pass


def lattice(nx, ny, box=(0.0, 1.0, 0.0, 1.0), jitter=1e-3, seed=0, bc=RIGID, flip_seed=0):
    """Jittered nx x ny lattice in box=(x0, x1, y0, y1)."""
    lib = _lib()
    p = _LatticeParams(nx, ny, box[0], box[1], box[2], box[3], jitter, seed, bc, flip_seed)
    h = lib.rm_lattice_build(ctypes.byref(p))
    if not h:
        raise ValueError("rm_lattice_build rejected the parameters")
    return _export(h, bc, lattice_ij=True)


_DELAUNAY_ERR = {-1: "bad input", -2: "no hull point found", -3: "a duplicate point, or one "
                 "inside the hull", -4: "flip budget spent", -5: "too many triangles",
                 -6: "a triangle turned clockwise", -7: "the hull is no longer convex"}


def delaunay(x, y):
    """(T, 3) counter-clockwise Delaunay triangles of the points (rm_delaunay);
    ValueError for input it cannot triangulate (e.g. duplicate points)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = len(x)
    tri = np.empty((2 * n, 3), dtype=np.int32)
    nt = ctypes.c_int32()
    rc = _lib().rm_delaunay(n, _ffi.ptr(x), _ffi.ptr(y), _ffi.ptr(tri), 2 * n, ctypes.byref(nt))
    if rc != 0:
        raise ValueError(f"rm_delaunay: {_DELAUNAY_ERR.get(rc, rc)}")
    return tri[:nt.value]


def points(x, y, box, margin=None):
    """Voronoi mesh of arbitrary points in the rigid-wall box (x0, x1, y0,
    y1), e.g. after the points moved.  Ghosts are the reflections across each
    wall of the points within `margin` of it (default: 3 mean spacings), as
    the lattice builder reflects its outer rows; rm_box_points_build
    triangulates points and ghosts and forms faces and geometry by the
    lattice builder's rules.  Cells keep the order of the points.
    ValueError for points it cannot triangulate (e.g. duplicates)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    N = len(x)
    x0, x1, y0, y1 = box
    if margin is None:
        margin = 3.0 * np.sqrt((x1 - x0) * (y1 - y0) / N)
    bp = _BoxPointsParams(N, _ffi.ptr(x), _ffi.ptr(y), x0, x1, y0, y1, margin)
    h = _lib().rm_box_points_build(ctypes.byref(bp))
    if not h:
        raise ValueError("rm_box_points_build could not triangulate the points")
    return _export(h, RIGID, lattice_ij=False)


class BoxMesher:
    """Meshes of n moving points in a rigid-wall box, one per step
    (rm_mover_mesh): the previous step's triangulation repaired by edge flips
    where it can be, built afresh otherwise; the same meshes as `points`."""

    def __init__(self, n, box, margin=None):
        x0, x1, y0, y1 = (float(v) for v in box)
        if margin is None:
            margin = 3.0 * np.sqrt((x1 - x0) * (y1 - y0) / n)
        self.n, self.box, self.margin = n, (x0, x1, y0, y1), margin
        p = _BoxPointsParams(n, None, None, x0, x1, y0, y1, margin)
        self._lib = _lib()
        self._h = self._lib.rm_mover_create(ctypes.byref(p))
        if not self._h:
            raise ValueError("rm_mover_create rejected the box")

    def mesh(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if x.shape != (self.n,) or y.shape != (self.n,):
            raise ValueError(f"expected {self.n} points")
        h = self._lib.rm_mover_mesh(self._h, _ffi.ptr(x), _ffi.ptr(y))
        if not h:
            raise ValueError("rm_mover_mesh could not triangulate the points")
        return _export(h, RIGID, lattice_ij=False)

    def permute(self, order):
        """Renumber the points, new point i = old point order[i] (rm_mover_permute):
        the kept triangulation follows, so the next mesh is repaired, not
        built afresh."""
        o = np.ascontiguousarray(order, dtype=np.int32)
        if o.shape != (self.n,) or self._lib.rm_mover_permute(self._h, _ffi.ptr(o)) != 0:
            raise ValueError("expected a permutation of the points")

    def counts(self):
        """(full triangulations, flip repairs) so far."""
        b, r = ctypes.c_int32(), ctypes.c_int32()
        self._lib.rm_mover_counts(self._h, ctypes.byref(b), ctypes.byref(r))
        return b.value, r.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rm_mover_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def from_triangles(x, y, ghost_orig, ghost_wall, tri, box):
    """The mesh of real points [0, N) and their wall reflections [N, N+G)
    from a triangulation the caller made (rm_points_build): x, y [N+G],
    ghost_orig and ghost_wall [G] (0 right, 1 up, 2 left, 3 down), tri
    (T, 3) point ids of the triangles with a real vertex."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    go = np.ascontiguousarray(ghost_orig, dtype=np.int32)
    gw = np.ascontiguousarray(ghost_wall, dtype=np.int32)
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    N = len(x) - len(go)
    x0, x1, y0, y1 = box
    p = _PointsParams(N, len(go), _ffi.ptr(x), _ffi.ptr(y), _ffi.ptr(go), _ffi.ptr(gw),
                      len(tri), _ffi.ptr(tri), x0, x1, y0, y1)
    h = _lib().rm_points_build(ctypes.byref(p))
    if not h:
        raise ValueError("rm_points_build rejected the input")
    return _export(h, RIGID, lattice_ij=False)


def _export(h, bc, lattice_ij):
    lib = _lib()
    handle = _Handle(h)
    n = [ctypes.c_int32() for _ in range(4)]
    lib.rm_sizes(h, *[ctypes.byref(x) for x in n])
    N, F, G, nnz = (x.value for x in n)
    i32 = lambda k: np.empty(k, dtype=np.int32)
    f64 = lambda k: np.empty(k, dtype=np.float64)
    arrs = dict(face_first=i32(F), face_second=i32(F), v1x=f64(F), v1y=f64(F), v2x=f64(F),
                v2y=f64(F), pt_x=f64(N + G), pt_y=f64(N + G), cm_x=f64(N + G),
                cm_y=f64(N + G), volume=f64(N), tess_volume=f64(N), csr_off=i32(N + 1),
                csr_face=i32(nnz), ghost_orig=i32(G),
                lat_i=i32(N) if lattice_ij else None, lat_j=i32(N) if lattice_ij else None)
    lib.rm_export(h, *[_ffi.ptr(arrs[k]) if arrs[k] is not None else None for k in (
        "face_first", "face_second", "v1x", "v1y", "v2x", "v2y", "pt_x", "pt_y", "cm_x", "cm_y",
        "volume", "tess_volume", "csr_off", "csr_face", "ghost_orig", "lat_i", "lat_j")])
    return Mesh(n_cells=N, n_faces=F, n_ghosts=G, bc=bc, _handle=handle, **arrs)
