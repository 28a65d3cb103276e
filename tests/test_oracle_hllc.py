"""Pins the oracle's EOS and HLLC (hllc.cpp:27-154, ideal_gas.cpp) against the
reference's own known-answer tests and against properties the mathematics
fixes.  Each check is one a plausible slip (a wrong sign in the frame shift,
swapped sides, a dropped star-state term) would fail."""
import json
import os

import numpy as np
import pytest

import pyoracle as po

G = 5.0 / 3.0
HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def prim(d, p, vx, vy, g=G):
    return (d, p, vx, vy, po.dp2e(g, d, p), po.dp2c(g, d, p))


def physical_flux(d, p, vx, vy, g=G):
    e = p / d / (g - 1)
    E = d * (e + 0.5 * (vx * vx + vy * vy))
    return np.array([d * vx, d * vx * vx + p, d * vx * vy, (E + p) * vx])


def test_eos_closed_forms():
    for d, p in [(1.0, 1.0), (0.125, 0.1), (2.0, 3e-5), (7.3, 1e4)]:
        e = po.dp2e(G, d, p)
        assert e == pytest.approx(p / (d * (G - 1)), rel=1e-15)
        assert po.dp2c(G, d, p) == pytest.approx(np.sqrt(G * p / d), rel=1e-15)


def test_hydrostatic_convection_mass_flux():
    """tests/newtonian/common/hydrostatic_convection_mass_flux: uniform state
    at rest, face velocity w: the mass flux is -rho*w (test.py compares the
    printed numbers for equality)."""
    k = KNOWN["hydrostatic_convection_mass_flux"]
    s = (k["density"], k["pressure"], 0.0, 0.0, k["energy"], k["sound_speed"])
    out = po.hllc(s, s, k["face_velocity"])
    assert out[0] == -k["density"] * k["face_velocity"]


def test_hll_acoustic_eulerian():
    """tests/newtonian/common/hll_acoustic_eulerian: a weak pressure jump;
    the flux matches the linearised star state within the grader's 1%."""
    k = KNOWN["hll_acoustic_eulerian"]
    L = prim(k["density"], k["pressure"] + k["dp"], 0.0, 0.0, k["gamma"])
    R = prim(k["density"], k["pressure"], 0.0, 0.0, k["gamma"])
    f = po.hllc(L, R, 0.0)
    dl, pl, vl, cl, el = L[0], L[1], L[2], L[5], L[4]
    dr, pr, vr, cr = R[0], R[1], R[2], R[5]
    ps = (cl * dl * pr + cr * dr * (pl + cl * dl * (vl - vr))) / (cl * dl + cr * dr)
    vs = ((-ps + pl) / (cl * dl) + (ps - pr) / (cr * dr) + vl + vr) / 2.0
    ds = dl + (ps - pl) / cl ** 2
    es = el + pl * (ds - dl) / dl ** 2
    exact = [vs * ds, ps + ds * vs ** 2, 0.0, vs * (ps + ds * es + 0.5 * ds * vs ** 2)]
    for num, ex in zip(f, exact):
        if not (num == 0 and ex == 0):
            assert abs(num - ex) / abs(num + ex) < k["threshold"]


def test_hll_advection_eulerian():
    """tests/newtonian/common/hll_advection_eulerian: a pure contact moving at
    v=4 -- the grader allows 1%; HLLC resolves contacts exactly, so the
    upwind physical flux is matched to round-off."""
    k = KNOWN["hll_advection_eulerian"]
    L = prim(k["left_density"], k["pressure"], k["velocity"], 0.0, k["gamma"])
    R = prim(k["right_density"], k["pressure"], k["velocity"], 0.0, k["gamma"])
    f = po.hllc(L, R, 0.0)
    exact = physical_flux(L[0], L[1], L[2], 0.0, k["gamma"])
    np.testing.assert_allclose(f, exact, rtol=1e-13, atol=1e-13)
    assert np.all(np.abs(f - exact) <= k["threshold"] * np.abs(f + exact) + 1e-300)


@pytest.mark.parametrize("seed", range(6))
def test_consistency_moving_face(seed):
    """Identical states give the exact flux through a face moving at w:
    F(U) - w U (checks the frame shift and its undoing, hllc.cpp:109-152)."""
    rng = np.random.default_rng(seed)
    d, p = rng.uniform(0.2, 3.0), rng.uniform(0.1, 5.0)
    vx, vy, w = rng.uniform(-2, 2, 3)
    s = prim(d, p, vx, vy)
    f = po.hllc(s, s, w)
    E = d * (s[4] + 0.5 * (vx * vx + vy * vy))
    U = np.array([d, d * vx, d * vy, E])
    np.testing.assert_allclose(f, physical_flux(d, p, vx, vy) - w * U, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_mirror_symmetry(seed):
    """Reflecting x -> -x swaps the sides: mass, y-momentum and energy fluxes
    change sign, x-momentum flux does not."""
    rng = np.random.default_rng(100 + seed)
    dl, dr = rng.uniform(0.2, 3.0, 2)
    pl, pr = rng.uniform(0.1, 5.0, 2)
    ul, ur, vl, vr, w = rng.uniform(-1.5, 1.5, 5)
    f = po.hllc(prim(dl, pl, ul, vl), prim(dr, pr, ur, vr), w)
    g = po.hllc(prim(dr, pr, -ur, vr), prim(dl, pl, -ul, vl), -w)
    np.testing.assert_allclose(g, [-f[0], f[1], -f[2], -f[3]], rtol=1e-11, atol=1e-12)


def test_supersonic_upwinding():
    """All waves to the right: the flux is the left physical flux exactly;
    all to the left: the right one."""
    L = prim(1.0, 1.0, 5.0, 0.3)
    R = prim(0.5, 0.7, 4.5, -0.2)
    np.testing.assert_array_equal(po.hllc(L, R, 0.0), po.hllc(L, L, 0.0))
    np.testing.assert_allclose(po.hllc(L, R, 0.0), physical_flux(1.0, 1.0, 5.0, 0.3), rtol=1e-14)
    L2 = prim(1.0, 1.0, -5.0, 0.3)
    R2 = prim(0.5, 0.7, -4.5, -0.2)
    np.testing.assert_allclose(po.hllc(L2, R2, 0.0), physical_flux(0.5, 0.7, -4.5, -0.2),
                               rtol=1e-14)


@pytest.mark.parametrize("seed", range(4))
def test_galilean_shift(seed):
    """Adding u0 to both states and to the face velocity leaves the mass flux
    unchanged and shifts the x-momentum flux by u0 times it."""
    rng = np.random.default_rng(200 + seed)
    dl, dr = rng.uniform(0.3, 2.0, 2)
    pl, pr = rng.uniform(0.3, 2.0, 2)
    ul, ur, vl, vr, w, u0 = rng.uniform(-1, 1, 6)
    f = po.hllc(prim(dl, pl, ul, vl), prim(dr, pr, ur, vr), w)
    g = po.hllc(prim(dl, pl, ul + u0, vl), prim(dr, pr, ur + u0, vr), w + u0)
    assert g[0] == pytest.approx(f[0], rel=1e-11, abs=1e-12)
    assert g[1] == pytest.approx(f[1] + u0 * f[0], rel=1e-11, abs=1e-12)


def test_nan_and_wave_speed_errors():
    ok = prim(1.0, 1.0, 0.0, 0.0)
    bad = (1.0, 1.0, float("nan"), 0.0, 1.5, 1.29)
    with pytest.raises(po.OracleError) as e:
        po.hllc(ok, bad, 0.0)
    assert po.CODES[e.value.code] == "Hllc::Solved entered with nan"
    with pytest.raises(po.OracleError) as e:
        po.hllc(bad, ok, 0.0)
    assert po.CODES[e.value.code] == "Invalid wave speeds in hllc solver"
