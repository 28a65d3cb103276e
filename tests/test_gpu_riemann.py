"""rg_riemann, the batched RiemannSolver (riemann_solver.hpp:22-23 ->
hllc.cpp:97-154): the reference's own HLLC known-answer tests run through
it, and every flux equals the oracle's Hllc bit for bit on random states in
all four wave patterns, with and without a moving face; its errors carry
the reference's messages and entries."""
import json
import os

import numpy as np
import pytest

import cases
import pyoracle as po

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))
G = 5.0 / 3.0


def _ctx():
    from rich_amd.gpu import RichGPU
    return RichGPU(dict(cases.small_meshes())["rigid"], order=2)


def prim(d, p, vx, vy, g=G):
    return [d, p, vx, vy, po.dp2e(g, d, p), po.dp2c(g, d, p)]


def oracle(L, R, W):
    """The oracle's Hllc per pair, in rg_riemann's order (m, E, mx, my)."""
    out = np.empty((4, L.shape[1]))
    for i in range(L.shape[1]):
        m, mx, my, E = po.hllc(tuple(L[:, i]), tuple(R[:, i]), float(W[i]))
        out[:, i] = (m, E, mx, my)
    return out


def test_reference_known_answers():
    g = _ctx()
    k = KNOWN["hydrostatic_convection_mass_flux"]
    s = np.array([[k["density"]], [k["pressure"]], [0.0], [0.0], [k["energy"]],
                  [k["sound_speed"]]])
    f = g.riemann(s, s, np.array([k["face_velocity"]]))
    assert f[0, 0] == -k["density"] * k["face_velocity"]
    k = KNOWN["hll_advection_eulerian"]
    L = np.array([prim(k["left_density"], k["pressure"], k["velocity"], 0.0, k["gamma"])]).T
    R = np.array([prim(k["right_density"], k["pressure"], k["velocity"], 0.0, k["gamma"])]).T
    f = g.riemann(L, R)
    d, p, v = k["left_density"], k["pressure"], k["velocity"]
    E = d * (p / d / (k["gamma"] - 1) + 0.5 * v * v)
    exact = [d * v, (E + p) * v, d * v * v + p, 0.0]
    assert np.all(np.abs(f[:, 0] - exact) <= k["threshold"] * np.abs(f[:, 0] + exact) + 1e-300)
    k = KNOWN["hll_acoustic_eulerian"]
    L = np.array([prim(k["density"], k["pressure"] + k["dp"], 0.0, 0.0, k["gamma"])]).T
    R = np.array([prim(k["density"], k["pressure"], 0.0, 0.0, k["gamma"])]).T
    np.testing.assert_array_equal(g.riemann(L, R), oracle(L, R, np.zeros(1)))
    g.close()


@pytest.mark.parametrize("moving", [False, True])
def test_random_states_equal_oracle(moving):
    rng = np.random.default_rng(7 if moving else 8)
    n = 4000
    d = rng.uniform(0.1, 3, (2, n))
    p = rng.uniform(0.1, 3, (2, n))
    v = rng.normal(0, 1.5, (2, 2, n)) * rng.choice([0.1, 1, 4], (1, 1, n))
    L = np.array([prim(d[0, i], p[0, i], v[0, 0, i], v[0, 1, i]) for i in range(n)]).T
    R = np.array([prim(d[1, i], p[1, i], v[1, 0, i], v[1, 1, i]) for i in range(n)]).T
    W = rng.normal(0, 2, n) if moving else np.zeros(n)
    g = _ctx()
    f = g.riemann(L, R, W if moving else None)
    ref = oracle(L, R, W)
    np.testing.assert_array_equal(f, ref)
    # every wave pattern is met: supersonic either way and both star regions
    ul, ur = L[2] - W, R[2] - W
    sl = np.minimum(ul - L[5], ur - R[5])
    sr = np.maximum(ul + L[5], ur + R[5])
    assert (sl > 0).any() and (sr < 0).any() and ((sl <= 0) & (sr >= 0)).sum() > n // 2
    g.close()


def test_errors():
    from rich_amd.gpu import RichGPUError
    g = _ctx()
    ok = np.array([prim(1.0, 1.0, 0.0, 0.0)] * 3).T
    bad = ok.copy()
    bad[2, 1] = np.nan
    with pytest.raises(RichGPUError) as e:
        g.riemann(ok, bad)
    assert e.value.text == "Hllc::Solved entered with nan" and e.value.index == 1
    assert e.value.entries == []
    W = np.array([0.0, 0.0, 0.5])
    with pytest.raises(RichGPUError) as e:
        g.riemann(bad[:, [0, 0, 1]], ok, W)
    assert e.value.text == "Invalid wave speeds in hllc solver" and e.value.index == 2
    names = [n for n, _ in e.value.entries]
    assert names[:2] == ["left density", "left pressure"] and names[-1] == "right wave speed"
    vals = np.array([v for _, v in e.value.entries])
    assert vals[0] == 1.0 and np.isnan(vals[2]) and vals[12] == 0.5
    np.testing.assert_array_equal(vals[6:12], [1.0, 1.0, -0.5, 0.0, ok[5, 0], ok[4, 0]])
    # a good call afterwards
    np.testing.assert_array_equal(g.riemann(ok, ok), oracle(ok, ok, np.zeros(3)))
    g.close()
