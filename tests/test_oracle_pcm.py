"""Pins the oracle's PCM reconstruction and ModularFluxCalculator over it
(interpolations/PCM.cpp:5-37, modular_flux_calculator.cpp:39-71).

* PCM's face states are the cells themselves, ghost sides the ghost
  generator's values: asserted exactly.
* On a Voronoi face the modular calculator's normal (vertex minus the
  first side's point, less its component along the face) points the same
  way as SimpleFluxCalculator's point difference, and at a rigid wall the
  ghost's reflection about the ghost-real axis equals Reflect about the
  face; rotate divides by |n|, so the two calculators' fluxes agree to
  round-off.  SimpleFluxCalculator is pinned by the first-order tests in
  test_oracle_scheme.py.
* First-order Sod against the exact Riemann solution.
"""
import numpy as np
import pytest

import cases
import exact_riemann
import pyoracle as po
from test_oracle_pins import ref_ghosts

G = 5.0 / 3.0


def _cells(m, seed):
    c = cases.rough_cells(m, seed=seed)
    c[0] = np.abs(c[0]) + 0.2
    c[1] = np.abs(c[1]) + 0.2
    c[2] += 0.3
    return c


@pytest.mark.parametrize("name", [n for n, _ in cases.small_meshes()] + ["merged"])
def test_pcm_face_states_are_the_cells(name):
    m = cases.merged_ghost_mesh()[0] if name == "merged" else dict(cases.small_meshes())[name]
    cells = _cells(m, 3)
    L, R = po.Oracle(m, order=3).pcm(cells)
    cng = ref_ghosts(m, cells)
    np.testing.assert_array_equal(L, cng[:, m.face_first])
    np.testing.assert_array_equal(R, cng[:, m.face_second])


@pytest.mark.parametrize("name", [n for n, _ in cases.small_meshes()])
def test_modular_pcm_equals_simple_flux(name):
    m = dict(cases.small_meshes())[name]
    cells = _cells(m, 5)
    f3 = po.Oracle(m, order=3).flux(cells)
    f1 = po.Oracle(m, order=1).flux(cells)
    # a face's direction is known to the vertices' round-off over its
    # length, and HLLC's cancellations amplify it: faces down to 1e-8 of a
    # spacing agree less closely than long ones
    L = np.hypot(m.v2x - m.v1x, m.v2y - m.v1y)
    tol = 1e-10 + 1e-12 * np.median(L) / L
    scale = np.maximum(np.abs(f1), 1e-3 * np.abs(f1).max(axis=1, keepdims=True))
    err = (np.abs(f3 - f1) / scale).max(axis=0)
    assert (err < tol).all(), (err / tol).max()
    # not the same function: the frames round differently
    assert not np.array_equal(f3, f1)


def test_sod_modular_pcm_against_exact_riemann():
    m = cases.sod_mesh(160, 10)
    cells = cases.sod_cells(m)
    o = po.Oracle(m, order=3)
    ext = o.init_extensives(cells)
    c, t = cells, 0.0
    while t < 0.1:
        c, ext, dts = o.run(c, ext, 1, integrator=1)
        t += dts[0]
    x = m.pt_x[:m.n_cells]
    d, u, p = exact_riemann.sample(x, t, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), G)
    assert np.mean(np.abs(c[0] - d)) < 0.035
    # rigid walls, first order: mass and energy conserved to round-off
    assert abs(ext[0].sum() / o.init_extensives(cells)[0].sum() - 1) < 1e-14
