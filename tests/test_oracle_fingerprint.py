"""The oracle's branch fingerprint (ro_fingerprint): one 64-bit sum per
step of the branches ro_run takes.  It must not depend on the thread count,
must not change where no branch changes, and must change where one does --
it decides where a central difference of the step may straddle a
discontinuity."""
import numpy as np

import cases
import pyoracle as po
from rich_amd import mesh as rmesh
from test_oracle_tangent import smooth_cells


def _run(m, cells, steps, **kw):
    o = po.Oracle(m, order=2, **kw)
    return o.run(cells, o.init_extensives(cells), steps, fingerprint=True)


def test_fingerprint_is_deterministic_and_thread_independent():
    m = dict(cases.small_meshes())["rigid_flip"]
    cells = cases.rough_cells(m, seed=3)
    cells[0] = np.abs(cells[0]) + 0.3
    cells[1] = np.abs(cells[1]) + 0.3
    a = _run(m, cells, 6)
    b = _run(m, cells, 6, nthreads=4)
    c = _run(m, cells, 6)
    np.testing.assert_array_equal(a[3], b[3])
    np.testing.assert_array_equal(a[3], c[3])
    # per step, one sum per kind: HLLC (2), CFL (3), the limiters (4, 5, 7), shock switch (6)
    assert a[3].shape == (6, 8)
    assert np.all(a[3][:, [2, 3, 6, 7]] != 0)
    assert len({tuple(r) for r in a[3].tolist()}) == 6
    # the record does not change the run
    o = po.Oracle(m, order=2)
    plain = o.run(cells, o.init_extensives(cells), 6)
    np.testing.assert_array_equal(plain[0], a[0])


def test_fingerprint_follows_branches():
    """On a smooth state a 1e-12 change of one interior cell keeps every
    branch; raising that cell's pressure threefold switches its shock
    detector and limiters: the first step's fingerprint changes."""
    m = rmesh.lattice(20, 16, (0.0, 1.0, 0.0, 0.8), jitter=3e-2, seed=2, bc=rmesh.PERIODIC)
    cells = smooth_cells(m)
    base = _run(m, cells, 4)[3]
    i = m.n_cells // 2
    tiny = cells.copy()
    tiny[0, i] *= 1 + 1e-12
    np.testing.assert_array_equal(_run(m, tiny, 4)[3], base)
    big = cells.copy()
    big[1, i] *= 3.0
    assert np.any(_run(m, big, 4)[3][0] != base[0])
    # ties in a uniform region: the first equal neighbour gives the
    # extremum; a change of 1e-12 in one of them makes it the extremum
    uni = cells.copy()
    uni[0] = 1.0
    fu = _run(m, uni, 1)[3]
    uni2 = uni.copy()
    uni2[0, i] = 1.0 + 1e-12
    assert np.any(_run(m, uni2, 1)[3][0] != fu[0])
