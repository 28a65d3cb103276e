"""The forward-mode build of the oracle (oracle/tangent.cpp: the same source
compiled with dual numbers), the CPU reference for the GPU adjoint.

Pins: (1) its values are the double oracle's bit for bit; (2) its tangents
are the derivatives of the step, checked against central finite differences
where the step is smooth (no slope limiter: LinearGaussImproved's limiter
gates switch psi discontinuously, e.g. from 1 to 0 at a local extremum, so
FD of a limited run does not converge to any derivative); (3) a planted
mistake in a tangent seed is caught by the FD check."""
import numpy as np
import pytest

import cases
import pyoracle as po

MESHES = dict(cases.small_meshes())


def smooth_cells(m):
    N = m.n_cells
    x, y = m.pt_x[:N], m.pt_y[:N]
    c = np.zeros((4, N))
    c[0] = 1 + 0.2 * np.sin(3 * x + 1) * np.cos(2 * y)
    c[1] = 1 + 0.1 * np.cos(2 * x + 3 * y)
    c[2] = 0.1 * np.sin(2 * y)
    c[3] = 0.1 * np.cos(3 * x)
    return c


@pytest.mark.parametrize("name", ["rigid", "periodic_flip"])
def test_tangent_values_equal_double_oracle(name):
    m = MESHES[name]
    cells = cases.rough_cells(m, seed=8)
    o = po.Oracle(m, order=2)
    d = np.random.default_rng(1).standard_normal((3, 4, m.n_cells))
    c1, _, x1, _, dt1 = o.run_tangent(cells, d, 4)
    c0, x0, dt0 = o.run(cells, o.init_extensives(cells), 4)
    np.testing.assert_array_equal(c1, c0)
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(dt1, dt0)


def _fd_error(o, cells, d, nsteps, dc, h=1e-6, integrator=2):
    errs = []
    for k in range(d.shape[0]):
        cp, _, _ = o.run(cells + h * d[k], o.init_extensives(cells + h * d[k]), nsteps,
                         integrator=integrator)
        cm, _, _ = o.run(cells - h * d[k], o.init_extensives(cells - h * d[k]), nsteps,
                         integrator=integrator)
        fd = (cp - cm) / (2 * h)
        errs.append(np.abs(fd - dc[k]).max() / np.abs(dc[k]).max())
    return max(errs)


@pytest.mark.parametrize("integrator", [1, 2])
@pytest.mark.parametrize("name", ["rigid", "periodic"])
def test_tangent_matches_finite_differences(name, integrator):
    m = MESHES[name]
    cells = smooth_cells(m)
    o = po.Oracle(m, order=2, slf=False)
    d = np.random.default_rng(2).standard_normal((2, 4, m.n_cells))
    _, dc, _, _, _ = o.run_tangent(cells, d, 3, integrator=integrator)
    err = _fd_error(o, cells, d, 3, dc, integrator=integrator)
    print(name, integrator, "max relative FD error", err)
    assert err < 1e-5


def test_tangent_first_order():
    m = cases.sod_mesh(32, 8)
    cells = smooth_cells(m)
    o = po.Oracle(m, order=1)
    d = np.random.default_rng(3).standard_normal((2, 4, m.n_cells))
    _, dc, _, _, _ = o.run_tangent(cells, d, 3, integrator=1)
    assert _fd_error(o, cells, d, 3, dc, integrator=1) < 1e-5


def test_planted_tangent_mistake_is_caught():
    """A tangent seeded along a different direction than the FD perturbs
    fails the same check."""
    m = MESHES["periodic"]
    cells = smooth_cells(m)
    o = po.Oracle(m, order=2, slf=False)
    d = np.random.default_rng(2).standard_normal((1, 4, m.n_cells))
    wrong = d.copy()
    wrong[0, 1] *= 1.01  # pressure component off by 1 %
    _, dc, _, _, _ = o.run_tangent(cells, wrong, 3)
    assert _fd_error(o, cells, d, 3, dc) > 1e-4
