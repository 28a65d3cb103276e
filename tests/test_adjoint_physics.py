"""The adjoint physics of the GPU step (huji-rich_amd/csrc/gpu/physics_adj.hpp,
host build in tests/adjcheck) against central differences of the forward
functions: HLLC in every wave pattern, primitive recovery, init_extensives
and the CFL inverse time step.  Each check is the identity
<w, J v> = <J^T w, v> with J v from central differences."""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "adjcheck", "build", "libadjcheck.so")
G = 5.0 / 3.0


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("tests/adjcheck not built")
    L = ctypes.CDLL(LIB)
    D = ctypes.POINTER(ctypes.c_double)
    L.ac_hllc.argtypes = [D, D, ctypes.c_double, D, D]
    L.ac_zero_face_speed.argtypes = [ctypes.c_double] * 3
    L.ac_zero_face_speed.restype = ctypes.c_double
    L.ac_hllc_vjp.argtypes = [D, D, ctypes.c_double, D, D, D, D]
    L.ac_recover_vjp.argtypes = [ctypes.c_double, ctypes.c_double, D, D, D]
    L.ac_init_vjp.argtypes = [ctypes.c_double, ctypes.c_double, D, D, D]
    L.ac_cfl_inv_vjp.argtypes = [ctypes.c_double, ctypes.c_double, D, ctypes.c_double, D]
    return L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _hllc(lib, L, R, fr):
    out = np.zeros(4)
    _, pl = _d(L)
    _, pr = _d(R)
    _, pf = _d(fr)
    rc = lib.ac_hllc(pl, pr, G, pf, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return out


def _frame(rng):
    nx, ny = rng.standard_normal(2)
    s = rng.uniform(0.5, 2.0)
    px, py = -ny * s, nx * s
    return np.array([nx, ny, np.hypot(nx, ny), px, py, np.hypot(px, py)])


def _branch(L, R, fr):
    an = fr[2]
    ul = (L[2] * fr[0] + L[3] * fr[1]) / an
    ur = (R[2] * fr[0] + R[3] * fr[1]) / an
    cl, cr = np.sqrt(G * L[1] / L[0]), np.sqrt(G * R[1] / R[0])
    sl, sr = min(ul - cl, ur - cr), max(ul + cl, ur + cr)
    ss = (R[1] - L[1] + L[0] * ul * (sl - ul) - R[0] * ur * (sr - ur)) / (
        L[0] * (sl - ul) - R[0] * (sr - ur))
    return 0 if sl > 0 else 1 if ss >= 0 else 2 if sr >= 0 else 3


@pytest.mark.parametrize("want", [0, 1, 2, 3])
def test_hllc_vjp_each_wave_pattern(lib, want):
    rng = np.random.default_rng(10 + want)
    done = 0
    for _ in range(4000):
        fr = _frame(rng)
        n = fr[:2] / fr[2]
        L = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), *rng.standard_normal(2)])
        R = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), *rng.standard_normal(2)])
        shift = {0: 3.0, 3: -3.0}.get(want, 0.0)
        L[2:] += shift * n
        R[2:] += shift * n
        if _branch(L, R, fr) != want:
            continue
        w = rng.standard_normal(4)
        dL, dR = rng.standard_normal(4), rng.standard_normal(4)
        h = 1e-6
        fd = (w @ _hllc(lib, L + h * dL, R + h * dR, fr) -
              w @ _hllc(lib, L - h * dL, R - h * dR, fr)) / (2 * h)
        Lb, Rb = np.zeros(4), np.zeros(4)
        _, pl = _d(L)
        _, pr = _d(R)
        _, pf = _d(fr)
        _, pw = _d(w)
        lib.ac_hllc_vjp(pl, pr, G, pf, pw, Lb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        Rb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        an = Lb @ dL + Rb @ dR
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1.0), (want, fd, an)
        done += 1
        if done == 25:
            break
    assert done == 25


def test_recover_vjp(lib):
    rng = np.random.default_rng(3)
    for _ in range(50):
        vol = rng.uniform(0.5, 2)
        c = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), *rng.standard_normal(2)])
        m = c[0] * vol
        x = np.array([m, c[1] / (G - 1) * vol + 0.5 * m * (c[2] ** 2 + c[3] ** 2), c[2] * m,
                      c[3] * m])

        def rec(x):
            d = x[0] / vol
            vx, vy = x[2] / x[0], x[3] / x[0]
            e = x[1] / x[0] - 0.5 * (vx * vx + vy * vy)
            return np.array([d, (G - 1) * e * d, vx, vy])

        w, dx = rng.standard_normal(4), rng.standard_normal(4)
        h = 1e-6
        fd = (w @ rec(x + h * dx) - w @ rec(x - h * dx)) / (2 * h)
        xb = np.zeros(4)
        _, px = _d(x)
        _, pw = _d(w)
        lib.ac_recover_vjp(G, vol, px, pw, xb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert abs(fd - xb @ dx) <= 1e-7 * max(1.0, abs(fd))


def test_init_vjp(lib):
    rng = np.random.default_rng(4)
    for _ in range(50):
        vol = rng.uniform(0.5, 2)
        c = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), *rng.standard_normal(2)])

        def init(c):
            m = vol * c[0]
            return np.array([m, c[1] / c[0] / (G - 1) * m + 0.5 * m * (c[2] ** 2 + c[3] ** 2),
                             c[2] * m, c[3] * m])

        w, dc = rng.standard_normal(4), rng.standard_normal(4)
        h = 1e-6
        fd = (w @ init(c + h * dc) - w @ init(c - h * dc)) / (2 * h)
        cb = np.zeros(4)
        _, pc = _d(c)
        _, pw = _d(w)
        lib.ac_init_vjp(G, vol, pc, pw, cb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert abs(fd - cb @ dc) <= 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("moving", [True, False])
def test_cfl_inv_vjp(lib, moving):
    rng = np.random.default_rng(5)
    for _ in range(50):
        width = rng.uniform(0.01, 0.1)
        c = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), 0.0, 0.0])
        if moving:
            c[2:] = rng.standard_normal(2)

        def inv(c):
            return 1.0 / ((np.sqrt(G * c[1] / c[0]) + np.hypot(c[2], c[3])) / width)

        dc = rng.standard_normal(4)
        if not moving:
            dc[2:] = 0.0  # |v| has no derivative at v = 0; only d, p vary
        h = 1e-6
        fd = (inv(c + h * dc) - inv(c - h * dc)) / (2 * h)
        cb = np.zeros(4)
        _, pc = _d(c)
        lib.ac_cfl_inv_vjp(G, width, pc, 1.0, cb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert abs(fd - cb @ dc) <= 1e-7 * max(1.0, abs(fd))


def test_zero_face_speed_is_the_quotient(lib):
    """The face kernels' face speed without face velocities skips the
    division where |n| > 0: bit for bit (n . 0) / |n| on signed zeros,
    infinities, NaNs, denormals and ordinary values."""
    vals = [0.0, -0.0, 5e-324, -5e-324, 1e-300, -1.0, 1.0, 3.7, -1e300, np.inf, -np.inf, np.nan]
    ans = [0.0, -0.0, 5e-324, 1e-300, 0.5, 1.0, 1e300, np.inf, -np.inf, -1.0, np.nan]
    n = 0
    with np.errstate(all="ignore"):
        for nx in vals:
            for ny in vals:
                for an in ans:
                    want = (np.float64(nx) * 0.0 + np.float64(ny) * 0.0) / np.float64(an)
                    got = lib.ac_zero_face_speed(nx, ny, an)
                    if np.isnan(want):
                        assert np.isnan(got), (nx, ny, an)
                    else:
                        assert np.float64(got).tobytes() == want.tobytes(), (nx, ny, an, got, want)
                    n += 1
    assert n == len(vals) ** 2 * len(ans)
