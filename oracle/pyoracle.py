"""ctypes binding of oracle/build/liboracle.so -- the CPU restatement of the
reference step, used only as the checker (tests, smoke) and as the timed CPU
baseline of bench.py.  Nothing in the product imports this module."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# RICH_ORACLE_LIB names another build of rich_oracle.c (the mutation test
# loads copies with planted mistakes)
LIB_PATH = os.environ.get("RICH_ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle.so")
TANGENT_PATH = os.path.join(_HERE, "build", "liboracle_tangent.so")

CODES = {
    0: "ok",
    1: "Negative density",
    2: "Negative thermal energy",
    3: "Hllc::Solved entered with nan",
    4: "Invalid wave speeds in hllc solver",
    5: "Singular matrix",
    6: "Cell has too many neighbors",
}


class Mesh_(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int32), ("n_faces", ctypes.c_int32),
                ("n_ghosts", ctypes.c_int32)] + [
        (k, ctypes.c_void_p) for k in (
            "face_first", "face_second", "v1x", "v1y", "v2x", "v2y", "pt_x", "pt_y", "cm_x",
            "cm_y", "volume", "tess_volume", "csr_off", "csr_face", "ghost_orig")]


class Params_(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
                ("order", ctypes.c_int32), ("bc", ctypes.c_int32), ("slf", ctypes.c_int32),
                ("delta_v", ctypes.c_double), ("theta", ctypes.c_double),
                ("delta_p", ctypes.c_double), ("nthreads", ctypes.c_int32)]


class Quad_(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("d", ctypes.c_void_p)]


class Error_(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("index", ctypes.c_int32)]


class OracleError(RuntimeError):
    def __init__(self, code, index):
        super().__init__(f"{CODES.get(code, code)} (index {index})")
        self.code = code
        self.index = index


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.ro_hllc.argtypes = [P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double,
                              P(ctypes.c_double)]
        L.ro_dp2e.restype = ctypes.c_double
        L.ro_dp2e.argtypes = [ctypes.c_double] * 3
        L.ro_dp2c.restype = ctypes.c_double
        L.ro_dp2c.argtypes = [ctypes.c_double] * 3
        L.ro_ghost_cells.argtypes = [P(Mesh_), P(Params_), Quad_]
        L.ro_lgi.restype = Error_
        L.ro_lgi.argtypes = [P(Mesh_), P(Params_), Quad_, ctypes.c_void_p, ctypes.c_void_p,
                             Quad_, Quad_]
        L.ro_pcm.argtypes = [P(Mesh_), P(Params_), Quad_, Quad_, Quad_]
        L.ro_fingerprint.restype = None
        L.ro_fingerprint.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        for n in ("ro_point_velocity", "ro_point_fix", "ro_face_velocity"):
            getattr(L, n).restype = None
        L.ro_point_velocity.argtypes = [P(Mesh_), ctypes.c_double] + [ctypes.c_void_p] * 4 + [
            ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.ro_point_fix.argtypes = [P(Mesh_), ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p]
        L.ro_face_velocity.argtypes = [P(Mesh_), ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.ro_flux.restype = Error_
        L.ro_flux.argtypes = [P(Mesh_), P(Params_), Quad_, ctypes.c_void_p, ctypes.c_void_p,
                              Quad_]
        L.ro_update_extensives.argtypes = [P(Mesh_), P(Params_), Quad_, ctypes.c_double, Quad_]
        L.ro_update_cells.restype = Error_
        L.ro_update_cells.argtypes = [P(Mesh_), P(Params_), Quad_, Quad_]
        L.ro_cfl.restype = ctypes.c_double
        L.ro_cfl.argtypes = [P(Mesh_), P(Params_), Quad_, ctypes.c_void_p, ctypes.c_void_p]
        L.ro_init_extensives.argtypes = [P(Mesh_), P(Params_), Quad_, Quad_]
        L.ro_run.restype = Error_
        L.ro_run.argtypes = [P(Mesh_), P(Params_), ctypes.c_int32, ctypes.c_int32,
                             ctypes.c_void_p, ctypes.c_void_p, Quad_, Quad_, ctypes.c_void_p]
        _lib = L
    return _lib


_tlib = None


def tlib():
    """The forward-mode build of the same source (dual numbers)."""
    global _tlib
    if _tlib is None:
        if not os.path.exists(TANGENT_PATH):
            raise RuntimeError(f"{TANGENT_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(TANGENT_PATH)
        P = ctypes.POINTER
        L.rt_k.restype = ctypes.c_int
        L.rt_run.restype = ctypes.c_int
        L.rt_run.argtypes = [P(Mesh_), P(Params_), ctypes.c_int32, ctypes.c_int32,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             P(ctypes.c_int32)]
        _tlib = L
    return _tlib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _quad(a):
    """(4, n) float64 array -> struct of 4 row pointers."""
    assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape[0] == 4
    return Quad_(*[ctypes.c_void_p(a[k].ctypes.data) for k in range(4)])


class Oracle:
    """Holds a mesh and scheme parameters; arrays are numpy (4, n) SoA:
    cells = (density, pressure, vx, vy), extensives/fluxes =
    (mass, energy, momentum x, momentum y)."""

    def __init__(self, mesh, gamma=5.0 / 3.0, cfl=0.3, order=2, slf=True, delta_v=0.2,
                 theta=0.5, delta_p=0.7, nthreads=1):
        self.mesh = mesh
        self._keep = [np.ascontiguousarray(getattr(mesh, k)) for k in (
            "face_first", "face_second", "v1x", "v1y", "v2x", "v2y", "pt_x", "pt_y", "cm_x",
            "cm_y", "volume", "tess_volume", "csr_off", "csr_face", "ghost_orig")]
        # an empty ghost map still gets a valid pointer; every array handed
        # to the oracle stays referenced by this object while it is in use
        gorig = self._keep[-1]
        self._one = np.zeros(1, np.int32)
        self.m = Mesh_(mesh.n_cells, mesh.n_faces, mesh.n_ghosts,
                       *[_p(a) if a.size else (_p(self._one) if a is gorig else None)
                         for a in self._keep])
        self.prm = Params_(gamma, cfl, order, mesh.bc, 1 if slf else 0, delta_v, theta,
                           delta_p, nthreads)
        self.N = mesh.n_cells
        self.NG = mesh.n_cells + mesh.n_ghosts
        self.F = mesh.n_faces

    def _cells_ng(self, cells):
        c = np.zeros((4, self.NG))
        c[:, :self.N] = cells[:, :self.N]
        return c

    @staticmethod
    def _check(e):
        if e.code != 0:
            raise OracleError(e.code, e.index)

    def ghost_cells(self, cells):
        c = self._cells_ng(cells)
        lib().ro_ghost_cells(ctypes.byref(self.m), ctypes.byref(self.prm), _quad(c))
        return c

    def lgi(self, cells):
        """LinearGaussImproved: (slopes (8,N), naive (8,N), left (4,F), right (4,F))."""
        c = self.ghost_cells(cells)
        s = np.zeros((8, self.N))
        nv = np.zeros((8, self.N))
        sp = (ctypes.c_void_p * 8)(*[s[k].ctypes.data for k in range(8)])
        np_ = (ctypes.c_void_p * 8)(*[nv[k].ctypes.data for k in range(8)])
        L = np.zeros((4, self.F))
        R = np.zeros((4, self.F))
        e = lib().ro_lgi(ctypes.byref(self.m), ctypes.byref(self.prm), _quad(c), sp, np_,
                         _quad(L), _quad(R))
        self._check(e)
        return s, nv, L, R

    def pcm(self, cells):
        """PCM: (left (4,F), right (4,F)) face states."""
        c = self._cells_ng(cells)
        L = np.zeros((4, self.F))
        R = np.zeros((4, self.F))
        lib().ro_pcm(ctypes.byref(self.m), ctypes.byref(self.prm), _quad(c), _quad(L), _quad(R))
        return L, R

    def point_velocity(self, cells, chi=0.15, eta=0.02):
        """Lagrangian + RoundCells point velocities (2, N)."""
        c = np.ascontiguousarray(cells[:, :self.N], dtype=np.float64)
        w = np.zeros((2, self.N))
        lib().ro_point_velocity(ctypes.byref(self.m), self.prm.gamma, *[_p(c[k]) for k in range(4)],
                                chi, eta, _p(w[0]), _p(w[1]))
        return w

    def point_fix(self, w, dt, box):
        """CorrectPointsOverShoot of the box (x0, x1, y0, y1): new (2, N)."""
        w = np.array(w, dtype=np.float64, order="C")
        b = np.ascontiguousarray(box, dtype=np.float64)
        lib().ro_point_fix(ctypes.byref(self.m), dt, _p(b), _p(w[0]), _p(w[1]))
        return w

    def face_velocity(self, w):
        """StationaryBox / CalcFaceVelocity face velocities (2, F)."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        fw = np.zeros((2, self.F))
        lib().ro_face_velocity(ctypes.byref(self.m), self.mesh.bc, _p(w[0]), _p(w[1]),
                               _p(fw[0]), _p(fw[1]))
        return fw

    def flux(self, cells, w=None):
        c = self._cells_ng(cells)
        fl = np.zeros((4, self.F))
        wx, wy = (None, None) if w is None else (np.ascontiguousarray(w[0]),
                                                 np.ascontiguousarray(w[1]))
        e = lib().ro_flux(ctypes.byref(self.m), ctypes.byref(self.prm), _quad(c), _p(wx), _p(wy),
                          _quad(fl))
        self._check(e)
        return fl

    def update_extensives(self, flux, dt, ext):
        ext = np.array(ext, dtype=np.float64, order="C")
        fl = np.ascontiguousarray(flux, dtype=np.float64)  # held through the call
        lib().ro_update_extensives(ctypes.byref(self.m), ctypes.byref(self.prm),
                                   _quad(fl), dt, _quad(ext))
        return ext

    def update_cells(self, ext):
        cells = np.zeros((4, self.N))
        x = np.ascontiguousarray(ext, dtype=np.float64)  # held through the call
        e = lib().ro_update_cells(ctypes.byref(self.m), ctypes.byref(self.prm),
                                  _quad(x), _quad(cells))
        self._check(e)
        return cells

    def cfl(self, cells, w=None):
        c = np.ascontiguousarray(cells)
        wx, wy = (None, None) if w is None else (np.ascontiguousarray(w[0]),
                                                 np.ascontiguousarray(w[1]))
        return lib().ro_cfl(ctypes.byref(self.m), ctypes.byref(self.prm), _quad(c), _p(wx), _p(wy))

    def init_extensives(self, cells):
        ext = np.zeros((4, self.N))
        c = np.ascontiguousarray(cells[:, :self.N], dtype=np.float64)  # held through the call
        lib().ro_init_extensives(ctypes.byref(self.m), ctypes.byref(self.prm),
                                 _quad(c), _quad(ext))
        return ext

    def run(self, cells, ext, nsteps, integrator=2, w=None, fingerprint=False):
        """hdsim TimeAdvance (1) / TimeAdvance2Heun (2); returns (cells, ext, dts),
        and with `fingerprint` also the branch fingerprints (nsteps, 8) uint64:
        per step, one sum per kind of branch (see ro_fingerprint)."""
        c = self._cells_ng(cells)
        x = np.array(ext, dtype=np.float64, order="C")
        dts = np.zeros(max(nsteps, 1))
        wx, wy = (None, None) if w is None else (np.ascontiguousarray(w[0]),
                                                 np.ascontiguousarray(w[1]))
        fp = np.zeros((max(nsteps, 1), 8), dtype=np.uint64) if fingerprint else None
        if fingerprint:
            lib().ro_fingerprint(_p(fp), nsteps)
        try:
            e = lib().ro_run(ctypes.byref(self.m), ctypes.byref(self.prm), integrator, nsteps,
                             _p(wx), _p(wy), _quad(c), _quad(x), _p(dts))
        finally:
            if fingerprint:
                lib().ro_fingerprint(None, 0)
        self._check(e)
        out = (c[:, :self.N].copy(), x, dts[:nsteps])
        return out + (fp[:nsteps],) if fingerprint else out

    def run_tangent(self, cells, dcells, nsteps, integrator=2, w=None):
        """Forward-mode derivatives of init_extensives + run: cells (4, N)
        initial state, dcells (K, 4, N) tangent directions of it (any K; run
        in batches of the build's width).  Returns (cells, dcells, ext,
        dext, dts) at the end: values as run() returns them (bit-identical)
        and their directional derivatives (K, 4, N)."""
        L = tlib()
        kb = L.rt_k()
        dcells = np.asarray(dcells, dtype=np.float64)
        K = dcells.shape[0]
        out_dc = np.zeros((K, 4, self.N))
        out_dx = np.zeros((K, 4, self.N))
        wx, wy = (None, None) if w is None else (np.ascontiguousarray(w[0]),
                                                 np.ascontiguousarray(w[1]))
        res = None
        for k0 in range(0, max(K, 1), kb):
            c = self._cells_ng(cells)
            dc = np.zeros((kb, 4, self.NG))
            nk = min(kb, K - k0)
            dc[:nk, :, :self.N] = dcells[k0:k0 + nk]
            x = np.zeros((4, self.N))
            dx = np.zeros((kb, 4, self.N))
            dts = np.zeros(max(nsteps, 1))
            idx = ctypes.c_int32(-1)
            rc = L.rt_run(ctypes.byref(self.m), ctypes.byref(self.prm), integrator, nsteps,
                          _p(wx), _p(wy), _p(c), _p(dc), _p(x), _p(dx), _p(dts),
                          ctypes.byref(idx))
            if rc != 0:
                raise OracleError(rc, idx.value)
            out_dc[k0:k0 + nk] = dc[:nk, :, :self.N]
            out_dx[k0:k0 + nk] = dx[:nk]
            res = (c[:, :self.N].copy(), x, dts[:nsteps])
        return res[0], out_dc, res[1], out_dx, res[2]


def hllc(left, right, velocity):
    """Hllc()(left, right, velocity); left/right = (d, p, vx, vy, e, c)."""
    L = (ctypes.c_double * 6)(*left)
    R = (ctypes.c_double * 6)(*right)
    out = (ctypes.c_double * 4)()
    rc = lib().ro_hllc(L, R, velocity, out)
    if rc != 0:
        raise OracleError(rc, 0)
    return np.array(out[:])


def dp2e(g, d, p):
    return lib().ro_dp2e(g, d, p)


def dp2c(g, d, p):
    return lib().ro_dp2c(g, d, p)
