"""The reference's acoustic test (tests/newtonian/two_dimensional/acoustic/
test.cpp, test.py, with its parameter files): a 60 x 60 Cartesian grid in
a periodic unit box, Eulerian (static mesh), LinearGaussImproved with
PeriodicGhostGenerator, HLLC through ModularFluxCalculator,
TimeAdvance2Heun, CFL 0.3, gamma 5/3; a sound wave of amplitude 1e-6 on
density 1 and pressure 0.6 (acoustic.cpp:7-16), run while t < 0.5.  The
grader requires the mean absolute difference from the wave translated by
c0 t to stay below 5e-8 for density, pressure and x velocity."""
import numpy as np

import pyoracle as po
from rich_amd import mesh as rmesh

GAMMA = 5.0 / 3.0
D0, P0, DD, T_END = 1.0, 0.6, 1e-6, 0.5
C0 = np.sqrt(GAMMA * P0 / D0)


def acoustic_mesh():
    return rmesh.lattice(60, 60, (0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=1, bc=rmesh.PERIODIC)


def wave(x):
    s = np.sin(2 * np.pi * x)
    return np.vstack([DD * s + D0, DD * C0 ** 2 * s + P0, C0 * DD / D0 * s, 0 * s])


def error_norms(m, cells, t):
    x = m.pt_x[:m.n_cells]
    want = wave(x - C0 * t)
    return [float(np.mean(np.abs(cells[k] - want[k]))) for k in range(3)]


def oracle_run(m):
    """Heun steps on the oracle while t < 0.5: (cells, t, dts)."""
    o = po.Oracle(m, order=2)
    c = wave(m.pt_x[:m.n_cells])
    x = o.init_extensives(c)
    t, dts = 0.0, []
    while t < T_END:
        c, x, dt = o.run(c, x, 1, integrator=2)
        t += float(dt[0])
        dts.append(float(dt[0]))
    return c, t, np.array(dts)


def test_acoustic_oracle():
    m = acoustic_mesh()
    cells, t, dts = oracle_run(m)
    err = error_norms(m, cells, t)
    print(f"{len(dts)} steps to t = {t:.4f}: errors {err}")
    assert len(dts) > 100
    assert all(e < 5e-8 for e in err)
    # the wave moved: the initial state is far off
    assert error_norms(m, wave(m.pt_x[:m.n_cells]), t)[0] > 5e-7
