"""The ensemble's disjoint-union mesh (rich_amd.ensemble.union): on the
union the oracle's slopes, face states and fluxes equal each instance's own,
concatenated, so renumbering cells, faces, ghosts and the CSR loses
nothing."""
import numpy as np
import pytest

import pyoracle as po
from rich_amd import ensemble
from rich_amd import mesh as rmesh


def _instances(bc):
    ms = [rmesh.lattice(n, n + 2, (-1.0, 1.0, -1.0, 1.0), jitter=2e-2, seed=10 + n, bc=bc,
                        flip_seed=n) for n in (6, 9, 7)]
    rng = np.random.default_rng(1)
    cells = []
    for m in ms:
        c = np.zeros((4, m.n_cells))
        c[0] = 1 + 0.3 * rng.random(m.n_cells)
        c[1] = 1 + 0.3 * rng.random(m.n_cells)
        c[2:] = 0.2 * rng.standard_normal((2, m.n_cells))
        cells.append(c)
    return ms, cells


@pytest.mark.parametrize("bc", [rmesh.RIGID, rmesh.PERIODIC])
def test_union_matches_instances(bc):
    ms, cells = _instances(bc)
    u = ensemble.union(ms)
    assert list(u.cell_off) == list(np.cumsum([0] + [m.n_cells for m in ms]))
    cu = np.concatenate(cells, axis=1)
    ou = po.Oracle(u.mesh, order=2)
    su, _, Lu, Ru = ou.lgi(cu)
    fu = ou.flux(cu)
    for k, (m, c) in enumerate(zip(ms, cells)):
        o = po.Oracle(m, order=2)
        s, _, L, R = o.lgi(c)
        f = o.flux(c)
        a, b = u.cell_off[k], u.cell_off[k + 1]
        fa, fb = u.face_off[k], u.face_off[k + 1]
        np.testing.assert_array_equal(su[:, a:b], s)
        np.testing.assert_array_equal(Lu[:, fa:fb], L)
        np.testing.assert_array_equal(Ru[:, fa:fb], R)
        np.testing.assert_array_equal(fu[:, fa:fb], f)


def test_sedov_ensemble_draws():
    ms, cells, prm = ensemble.sedov_ensemble(K=6, n_min=8, n_max=16, seed=3)
    assert len(ms) == 6 and all(8 <= n <= 16 for n in prm["n"])
    assert np.all((prm["energy"] >= 0.1) & (prm["energy"] <= 10))
    for m, c, e in zip(ms, cells, prm["energy"]):
        hot = c[1] > 1e-5
        assert np.isclose(np.sum(c[1][hot] * m.volume[hot]) / (2.0 / 3.0), e)
