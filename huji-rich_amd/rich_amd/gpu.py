"""ctypes binding of lib/librichgpu.so (include/richgpu.h): the hand-written
HIP step for gfx950.  There is no CPU fallback: every method calls the
native library, and constructing a context without a GPU raises."""
import ctypes

import numpy as np

from . import _ffi

RG_FP64 = 0
RG_MIXED = 1


class Params(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
                ("order", ctypes.c_int32), ("bc", ctypes.c_int32), ("slf", ctypes.c_int32),
                ("delta_v", ctypes.c_double), ("theta", ctypes.c_double),
                ("delta_p", ctypes.c_double), ("precision", ctypes.c_int32)]


class MeshS(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int32), ("n_faces", ctypes.c_int32),
                ("n_ghosts", ctypes.c_int32)] + [
        (k, ctypes.c_void_p) for k in (
            "face_first", "face_second", "v1x", "v1y", "v2x", "v2y", "pt_x", "pt_y", "cm_x",
            "cm_y", "volume", "tess_volume", "csr_off", "csr_face", "ghost_orig")]


class ErrorEntry(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("value", ctypes.c_double)]


class Error(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("index", ctypes.c_int32), ("step", ctypes.c_int32),
                ("phase", ctypes.c_int32), ("message", ctypes.c_char * 96),
                ("n_entries", ctypes.c_int32), ("entries", ErrorEntry * 16)]


class RichGPUError(RuntimeError):
    """A device-detected error, carrying the reference's message text."""

    def __init__(self, status, err):
        self.status = status
        self.code = err.code
        self.index = err.index
        self.step = err.step
        self.phase = err.phase
        self.text = err.message.decode(errors="replace")
        # the reference's UniversalError entries, in its order
        self.entries = [(err.entries[k].name.decode(errors="replace"), err.entries[k].value)
                        for k in range(max(0, min(err.n_entries, 16)))]
        super().__init__(f"{self.text} (code {self.code}, index {self.index}, step {self.step}, "
                         f"phase {self.phase}, status {status})")


MESH_FIELDS = ("face_first", "face_second", "v1x", "v1y", "v2x", "v2y", "pt_x", "pt_y", "cm_x",
               "cm_y", "volume", "tess_volume", "csr_off", "csr_face", "ghost_orig")


def lib():
    L = _ffi.load("librichgpu")
    if getattr(L, "_typed", False):
        return L
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    L.rg_create.argtypes = [P(Params), P(vp)]
    L.rg_destroy.argtypes = [vp]
    L.rg_destroy.restype = None
    L.rg_set_stream.argtypes = [vp, vp]
    L.rg_set_mesh.argtypes = [vp, P(MeshS)]
    L.rg_set_face_velocity.argtypes = [vp, vp, vp]
    L.rg_set_box.argtypes = [vp, vp]
    L.rg_get_geometry.argtypes = [vp, vp, vp, vp, vp]
    L.rg_replace_mesh.argtypes = [vp, P(MeshS)]
    L.rg_permute_cells.argtypes = [vp, vp]
    L.rg_moving_begin.argtypes = [vp, ctypes.c_double, ctypes.c_double, vp,
                                  P(ctypes.c_double), vp, vp]
    L.rg_moving_end.argtypes = [vp]
    L.rg_riemann.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp]
    for n in ("rg_set_cells", "rg_get_cells", "rg_set_extensives", "rg_get_extensives",
              "rg_set_fluxes", "rg_flux"):
        getattr(L, n).argtypes = [vp, vp, vp, vp, vp]
    L.rg_init_extensives.argtypes = [vp]
    L.rg_time_step.argtypes = [vp, P(ctypes.c_double)]
    L.rg_reconstruct.argtypes = [vp, vp, vp, vp]
    L.rg_update_extensives.argtypes = [vp, ctypes.c_double]
    L.rg_update_cells.argtypes = [vp]
    L.rg_advance.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
    L.rg_advance_noalloc.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
    L.rg_adjoint_forward.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
    L.rg_set_instances.argtypes = [vp, ctypes.c_int32, vp]
    L.rg_ensemble_config.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    L.rg_ensemble_autotune.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32,
                                       P(ctypes.c_int32)]
    L.rg_adjoint_backward.argtypes = [vp, vp, vp, vp]
    L.rg_last_error.argtypes = [vp, P(Error)]
    L.rg_profile_enable.argtypes = [vp, ctypes.c_int32]
    L.rg_kernel_count.restype = ctypes.c_int32
    L.rg_kernel_count.argtypes = []
    L.rg_kernel_name.restype = ctypes.c_char_p
    L.rg_kernel_name.argtypes = [ctypes.c_int32]
    L.rg_profile_read.argtypes = [vp, ctypes.c_int32, P(ctypes.c_int64), P(ctypes.c_double)]
    L.rg_device_bytes.restype = ctypes.c_int64
    L.rg_device_bytes.argtypes = [vp]
    L.rg_tiles.restype = ctypes.c_int32
    L.rg_tiles.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int32)]
    L.rg_use_tiles.argtypes = [vp, ctypes.c_int32]
    L._typed = True
    return L


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class RichGPU:
    """One device context: a mesh, the scheme parameters and the state.

    Arrays are numpy SoA: cells (4, N) = density, pressure, vx, vy;
    extensives and fluxes (4, n) = mass, energy, momentum x, momentum y;
    slopes (8, N) = d/dx of the four cell fields, then d/dy."""

    GEOMETRY = ("cm_x", "cm_y", "volume", "tess_volume")

    def __init__(self, mesh, gamma=5.0 / 3.0, cfl=0.3, order=2, slf=True, delta_v=0.2,
                 theta=0.5, delta_p=0.7, precision=RG_FP64, device_geometry=False, box=None):
        """device_geometry: upload no volumes or centres of mass and let the
        device form them (a rigid mesh then needs its box (x0, x1, y0, y1))."""
        self._lib = lib()
        self.mesh = mesh
        self.N, self.F = mesh.n_cells, mesh.n_faces
        prm = Params(gamma, cfl, order, mesh.bc, 1 if slf else 0, delta_v, theta, delta_p,
                     precision)
        h = ctypes.c_void_p()
        self._check(self._lib.rg_create(ctypes.byref(prm), ctypes.byref(h)), handle=None)
        self._h = h
        self.device_geometry = device_geometry
        if box is not None:
            self.set_box(box)
        self._upload(self._lib.rg_set_mesh, mesh)

    def set_box(self, box):
        b = np.ascontiguousarray(box, dtype=np.float64)
        if b.shape != (4,):
            raise ValueError("box is (x0, x1, y0, y1)")
        self._check(self._lib.rg_set_box(self._h, _p(b)))

    def geometry(self):
        """(volume, tess_volume, cm_x, cm_y) as the device holds them."""
        N, NG = self.N, self.N + self.mesh.n_ghosts
        v, tv, cx, cy = np.empty(N), np.empty(N), np.empty(NG), np.empty(NG)
        self._check(self._lib.rg_get_geometry(self._h, _p(v), _p(tv), _p(cx), _p(cy)))
        return v, tv, cx, cy

    def _upload(self, fn, mesh):
        keep = []
        ptrs = []
        for k in MESH_FIELDS:
            if self.device_geometry and k in self.GEOMETRY:
                ptrs.append(None)
                continue
            a = np.ascontiguousarray(getattr(mesh, k))
            if a.size == 0:
                a = np.zeros(1, dtype=a.dtype)
            keep.append(a)
            ptrs.append(_p(a))
        ms = MeshS(mesh.n_cells, mesh.n_faces, mesh.n_ghosts, *ptrs)
        self._check(fn(self._h, ctypes.byref(ms)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, handle=0):
        if rc == 0:
            return
        h = self._h if handle == 0 else handle
        err = Error()
        if h:
            self._lib.rg_last_error(h, ctypes.byref(err))
        raise RichGPUError(rc, err)

    @staticmethod
    def _rows(a, n, k=4):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (k, n):
            raise ValueError(f"expected shape {(k, n)}, got {a.shape}")
        return a

    def set_face_velocity(self, w):
        if w is None:
            self._check(self._lib.rg_set_face_velocity(self._h, None, None))
            return
        self._wkeep = self._rows(w, self.F, 2)
        self._check(self._lib.rg_set_face_velocity(self._h, _p(self._wkeep[0]),
                                                   _p(self._wkeep[1])))

    def moving_begin(self, chi, eta, box, velocities=True):
        """Stage 1 of a moving-mesh Heun step: (dt, vx, vy) with the point
        velocities (None, None without `velocities`)."""
        b = np.ascontiguousarray(box, dtype=np.float64)
        if b.shape != (4,):
            raise ValueError("box is (x0, x1, y0, y1)")
        dt = ctypes.c_double()
        vx = np.empty(self.N) if velocities else None
        vy = np.empty(self.N) if velocities else None
        self._check(self._lib.rg_moving_begin(self._h, float(chi), float(eta), _p(b),
                                              ctypes.byref(dt), _p(vx), _p(vy)))
        return dt.value, vx, vy

    def replace_mesh(self, mesh):
        """The rebuilt mesh of the moved points; the device state is kept."""
        if mesh.n_cells != self.N:
            raise ValueError("replace_mesh keeps the cell count")
        self._upload(self._lib.rg_replace_mesh, mesh)
        self.mesh = mesh
        self.F = mesh.n_faces

    def permute_cells(self, order):
        """New cell i = old cell order[i]: cells, extensives, mid extensives
        and point velocities follow (then replace_mesh with the mesh of the
        reordered points)."""
        o = np.ascontiguousarray(order, dtype=np.int32)
        if o.shape != (self.N,):
            raise ValueError(f"expected {self.N} indices")
        self._check(self._lib.rg_permute_cells(self._h, _p(o)))

    def moving_end(self):
        """Stage 2 of the moving-mesh step on the replaced mesh."""
        self._check(self._lib.rg_moving_end(self._h))

    def set_cells(self, cells):
        c = self._rows(cells, self.N)
        self._check(self._lib.rg_set_cells(self._h, *[_p(c[k]) for k in range(4)]))

    def get_cells(self):
        c = np.empty((4, self.N))
        self._check(self._lib.rg_get_cells(self._h, *[_p(c[k]) for k in range(4)]))
        return c

    def set_extensives(self, ext):
        e = self._rows(ext, self.N)
        self._check(self._lib.rg_set_extensives(self._h, *[_p(e[k]) for k in range(4)]))

    def get_extensives(self):
        e = np.empty((4, self.N))
        self._check(self._lib.rg_get_extensives(self._h, *[_p(e[k]) for k in range(4)]))
        return e

    def set_fluxes(self, flux):
        f = self._rows(flux, self.F)
        self._check(self._lib.rg_set_fluxes(self._h, *[_p(f[k]) for k in range(4)]))

    def init_extensives(self):
        self._check(self._lib.rg_init_extensives(self._h))

    def time_step(self):
        dt = ctypes.c_double()
        self._check(self._lib.rg_time_step(self._h, ctypes.byref(dt)))
        return dt.value

    def reconstruct(self):
        s = np.empty((8, self.N))
        L = np.empty((4, self.F))
        R = np.empty((4, self.F))
        self._check(self._lib.rg_reconstruct(self._h, _p(s), _p(L), _p(R)))
        return s, L, R

    def flux(self):
        f = np.empty((4, self.F))
        self._check(self._lib.rg_flux(self._h, *[_p(f[k]) for k in range(4)]))
        return f

    def riemann(self, left, right, velocity=None):
        """rg_riemann: HLLC on n pairs of face-frame Primitives, left and
        right (6, n) (density, pressure, x and y velocity, energy, sound
        speed), face velocities (n,) or None; returns the fluxes (4, n)
        (mass, energy, x momentum, y momentum)."""
        L = np.ascontiguousarray(left, dtype=np.float64)
        R = np.ascontiguousarray(right, dtype=np.float64)
        if L.ndim != 2 or L.shape[0] != 6 or R.shape != L.shape:
            raise ValueError("left and right are (6, n)")
        n = L.shape[1]
        W = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64)
        if W is not None and W.shape != (n,):
            raise ValueError("velocity is (n,)")
        out = np.empty((4, n))
        self._check(self._lib.rg_riemann(self._h, n, _p(L), _p(R), _p(W), _p(out)))
        return out

    def update_extensives(self, dt):
        self._check(self._lib.rg_update_extensives(self._h, float(dt)))

    def update_cells(self):
        self._check(self._lib.rg_update_cells(self._h))

    def advance(self, nsteps, integrator=2, record_dt=True):
        dts = np.empty(max(nsteps, 1)) if record_dt else None
        self._check(self._lib.rg_advance(self._h, integrator, nsteps, _p(dts)))
        return dts[:nsteps] if record_dt else None

    def advance_noalloc(self, nsteps, integrator=2, record_dt=True):
        """rg_advance_noalloc: rg_advance with no device allocation in the
        call (scratch in the library's fixed device array)."""
        dts = np.empty(max(nsteps, 1)) if record_dt else None
        self._check(self._lib.rg_advance_noalloc(self._h, integrator, nsteps, _p(dts)))
        return dts[:nsteps] if record_dt else None

    def adjoint_forward(self, nsteps, every=20, record_dt=True):
        """rg_advance(nsteps, Heun) keeping device checkpoints every `every`
        steps for adjoint_backward; needs init_extensives() of the cells."""
        dts = np.empty(max(nsteps, 1)) if record_dt else None
        # the device keeps one recorded run: a later forward replaces it
        self.adjoint_generation = getattr(self, "adjoint_generation", 0) + 1
        self._check(self._lib.rg_adjoint_forward(self._h, nsteps, every, _p(dts)))
        return dts[:nsteps] if record_dt else None

    def adjoint_backward(self, gcells, gext=None):
        """d(loss)/d(initial cells) (4, N) from d(loss)/d(final cells) (4, N)
        (and optionally d(loss)/d(final extensives))."""
        g = self._rows(gcells, self.N)
        x = None if gext is None else self._rows(gext, self.N)
        out = np.empty((4, self.N))
        self._check(self._lib.rg_adjoint_backward(self._h, _p(g), _p(x), _p(out)))
        return out

    def set_instances(self, cell_off):
        """Ensemble mode: the mesh is a disjoint union; cell_off (K+1,)."""
        self._inst_off = np.ascontiguousarray(cell_off, dtype=np.int32)
        self.K = len(self._inst_off) - 1
        self._check(self._lib.rg_set_instances(self._h, self.K, _p(self._inst_off)))

    def ensemble_config(self, block, blocks_per_cu):
        self._check(self._lib.rg_ensemble_config(self._h, block, blocks_per_cu))

    def ensemble_autotune(self, trial_steps=5):
        """[(block, blocks_per_cu, ms per step)] of every launch shape tried;
        the fastest is kept."""
        table = np.zeros((32, 3))
        rows = ctypes.c_int32()
        self._check(self._lib.rg_ensemble_autotune(self._h, trial_steps, _p(table), 32,
                                                   ctypes.byref(rows)))
        return [(int(b), int(p), float(ms)) for b, p, ms in table[:rows.value]]

    def advance_ensemble(self, nsteps, record_dt=True):
        """rg_advance in ensemble mode; dts (nsteps, K)."""
        dts = np.empty((max(nsteps, 1), self.K)) if record_dt else None
        self._check(self._lib.rg_advance(self._h, 2, nsteps, _p(dts)))
        return dts[:nsteps] if record_dt else None

    def set_stream(self, stream_ptr):
        self._check(self._lib.rg_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    def profile(self, on=True):
        self._check(self._lib.rg_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        """{kernel name: (launches, total ms)} since profile(True)."""
        out = {}
        for k in range(self._lib.rg_kernel_count()):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            self._check(self._lib.rg_profile_read(self._h, k, ctypes.byref(n), ctypes.byref(ms)))
            if n.value:
                out[self._lib.rg_kernel_name(k).decode()] = (n.value, ms.value)
        return out

    def device_bytes(self):
        return self._lib.rg_device_bytes(self._h)

    def use_tiles(self, on=True):
        """Allow (True, the default) or forbid the fused tile stages."""
        self._check(self._lib.rg_use_tiles(self._h, 1 if on else 0))

    def tiles(self):
        """(fused, face entries per stage, largest tile's entries): whether the
        second-order Heun loops run their stages fused per tile of 256 cells."""
        n = ctypes.c_int64()
        mx = ctypes.c_int32()
        on = self._lib.rg_tiles(self._h, ctypes.byref(n), ctypes.byref(mx))
        return bool(on), n.value, mx.value
