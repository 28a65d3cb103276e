"""Device-detected errors carry the reference's UniversalError: message,
first index and the entries the reference adds (simple_cell_updater.cpp:
50-57, hllc.cpp:46-71), with their names, order and values; and a step
that fails leaves the state as the failing call read it."""
import numpy as np
import pytest

import cases
import pyoracle as po
from parity import assert_parity

pytestmark = pytest.mark.gpu

G = 5.0 / 3.0
WAVE_NAMES = ["left density", "left pressure", "left x velocity", "left y velocity",
              "left sound speed", "left energy", "right density", "right pressure",
              "right x velocity", "right y velocity", "right sound speed", "right energy",
              "interface velocity", "left wave speed", "center wave speed", "right wave speed"]
CELL_NAMES = ["Cell index", "Cell mass", "Cell x momentum", "Cell y momentum", "Cell energy"]


def _mesh():
    return dict(cases.small_meshes())["rigid"]


def _cells(m, seed=8):
    c = cases.rough_cells(m, seed=seed)
    c[0] = np.abs(c[0]) + 0.5
    c[1] = np.abs(c[1]) + 0.5
    return c


def test_negative_thermal_energy_entries():
    from rich_amd.gpu import RichGPU, RichGPUError
    m = _mesh()
    cells = _cells(m)
    o = po.Oracle(m, order=2)
    bad = o.init_extensives(cells)
    bad[1, [17, 40]] = 0.0
    g = RichGPU(m, order=2)
    g.set_extensives(bad)
    with pytest.raises(RichGPUError) as e:
        g.update_cells()
    assert e.value.text == "Negative thermal energy" and e.value.index == 17
    assert [n for n, _ in e.value.entries] == CELL_NAMES
    vals = [v for _, v in e.value.entries]
    assert vals == [17.0, bad[0, 17], bad[2, 17], bad[3, 17], bad[1, 17]]
    g.close()


@pytest.mark.parametrize("order", [1, 2, 3])
def test_invalid_wave_speed_entries(order):
    """A NaN pressure reaches Hllc: its wave speeds fail every test of
    hllc.cpp:134-148; the entries are the rotated, frame-shifted states Hllc
    received and its wave speeds."""
    from rich_amd.gpu import RichGPU, RichGPUError
    m = _mesh()
    cells = _cells(m, seed=9)
    cells[1, 33] = np.nan
    o = po.Oracle(m, order=order)
    with pytest.raises(po.OracleError) as eo:
        o.flux(cells)
    g = RichGPU(m, order=order)
    g.set_cells(cells)
    with pytest.raises(RichGPUError) as e:
        g.flux()
    f = e.value.index
    assert e.value.text == "Invalid wave speeds in hllc solver" == po.CODES[eo.value.code]
    assert f == eo.value.index
    assert [n for n, _ in e.value.entries] == WAVE_NAMES
    got = np.array([v for _, v in e.value.entries])
    # the face states the flux calculator solved, from the oracle
    if order == 2:
        _, _, L, R = o.lgi(cells)
        left, right = L[:, f], R[:, f]
    else:
        L, R = po.Oracle(m, order=3).pcm(cells)
        left, right = L[:, f], R[:, f]
    a, b = m.face_first[f], m.face_second[f]
    px, py = m.v2x[f] - m.v1x[f], m.v2y[f] - m.v1y[f]
    if order == 1:
        nx, ny = m.pt_x[b] - m.pt_x[a], m.pt_y[b] - m.pt_y[a]
    else:  # calc_parallel_normal (modular_flux_calculator.cpp:12-22)
        if a < m.n_cells:
            qx, qy = m.v1x[f] - m.pt_x[a], m.v1y[f] - m.pt_y[a]
        else:
            qx, qy = m.pt_x[b] - m.v1x[f], m.pt_y[b] - m.v1y[f]
        sp = (qx * px + qy * py) / (px * px + py * py)
        nx, ny = qx - px * sp, qy - py * sp
    an, ap = np.hypot(nx, ny), np.hypot(px, py)
    want = []
    us, cs = [], []
    for s in (left, right):
        u = (s[2] * nx + s[3] * ny) / an
        c = np.sqrt(G * s[1] / s[0])
        want += [s[0], s[1], u, (s[2] * px + s[3] * py) / ap, c, s[1] / s[0] / (G - 1)]
        us.append(u)
        cs.append(c)
    sl = min(us[0] - cs[0], us[1] - cs[1])
    sr = max(us[0] + cs[0], us[1] + cs[1])
    with np.errstate(invalid="ignore"):
        ss = (right[1] - left[1] + left[0] * us[0] * (sl - us[0]) - right[0] * us[1] * (sr - us[1])) \
            / (left[0] * (sl - us[0]) - right[0] * (sr - us[1]))
    want += [0.0, sl, ss, sr]
    want = np.array(want)
    # NaN where the state carried the NaN, the same finite values elsewhere
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    assert np.isnan(got).any()
    # rotated velocities: the oracle's states rotate in a different operation order here
    assert_parity(got[None, :], want[None, :], rtol=1e-13, what="entries")
    g.close()


@pytest.mark.parametrize("nsteps", [1, 2, 3])
def test_failed_step_keeps_the_state_and_reports_entries(nsteps):
    """rg_advance with a cell whose energy cannot hold its kinetic energy:
    the stage-1 cell update of step 0 fails; the error names step 0, phase
    3 and carries that cell's mid-stage extensives (the oracle's), and the
    context's cells and extensives are the ones the step started from.  The
    fused step loop alternates two state buffers, so odd and even step
    counts end on different buffers."""
    from rich_amd.gpu import RichGPU, RichGPUError
    m = _mesh()
    cells = _cells(m, seed=10)
    cells[2, 55] = 0.8
    o = po.Oracle(m, order=2)
    ext = o.init_extensives(cells)
    ext[1, 55] = 0.1 * ext[0, 55] * (cells[2, 55] ** 2 + cells[3, 55] ** 2)
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    g.set_extensives(ext)
    with pytest.raises(RichGPUError) as e:
        g.advance(nsteps, integrator=2)
    assert e.value.text == "Negative thermal energy"
    assert (e.value.step, e.value.phase, e.value.index) == (0, 3, 55)
    dt = o.cfl(cells)
    mid = ext.copy()
    mid[0] += dt * 0.0
    mid[2] += 0.0 * dt
    mid[3] += 0.0 * dt
    mid[1] += dt * 0.0
    mid = o.update_extensives(o.flux(cells), dt, mid)
    assert [n for n, _ in e.value.entries] == CELL_NAMES
    assert [v for _, v in e.value.entries] == [55.0, mid[0, 55], mid[2, 55], mid[3, 55],
                                               mid[1, 55]]
    np.testing.assert_array_equal(g.get_cells(), cells)
    np.testing.assert_array_equal(g.get_extensives(), ext)
    g.close()


@pytest.mark.parametrize("k", [20, 21])
def test_neighbour_limit(k):
    """calc_naive_slope refuses a cell with more than 20 faces
    (LinearGaussImproved.cpp:72-79): "Cell has too many neighbors" with the
    cell's point; 20 faces are accepted, with the oracle's slopes."""
    from rich_amd.gpu import RichGPU, RichGPUError
    m = cases.wheel_mesh(k)
    N = m.n_cells
    cells = cases.rough_cells(m, seed=4)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    if k <= 20:
        s, _, L, R = po.Oracle(m, order=2).lgi(cells)
        gs, gl, gr = g.reconstruct()
        np.testing.assert_array_equal(gs, s)
        np.testing.assert_array_equal(gl, L)
        g.close()
        return
    for call in (g.reconstruct, g.flux):
        with pytest.raises(RichGPUError) as e:
            call()
        assert e.value.text == "Cell has too many neighbors" and e.value.index == 0
        assert e.value.entries == [("Cell x cor", m.pt_x[0]), ("Cell y cor", m.pt_y[0])]
    g.close()
    assert N > k


def test_singular_matrix_entries():
    """A gradient matrix singular to rounding (LinearGaussImproved.cpp:105-118):
    "Singular matrix" with the cell's point, its volume and the determinant,
    as the oracle fails on the same cell."""
    from rich_amd.gpu import RichGPU, RichGPUError
    m, i, lam = cases.singular_cell_mesh()
    cells = cases.rough_cells(m, seed=2)
    cells[0] = np.abs(cells[0]) + 0.5
    cells[1] = np.abs(cells[1]) + 0.5
    with pytest.raises(po.OracleError) as eo:
        po.Oracle(m, order=2).lgi(cells)
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    with pytest.raises(RichGPUError) as e:
        g.reconstruct()
    assert e.value.text == "Singular matrix" == po.CODES[eo.value.code]
    assert e.value.index == i == eo.value.index
    names = [n for n, _ in e.value.entries]
    vals = [v for _, v in e.value.entries]
    assert names == ["Cell x cor", "Cell y cor", "Cell volume", "Det was"]
    assert vals[:3] == [m.pt_x[i], m.pt_y[i], lam]
    assert abs(vals[3]) < 1e-10 * lam * lam
    g.close()
