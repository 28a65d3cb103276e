"""The two errors of LinearGaussImproved::calc_naive_slope
(LinearGaussImproved.cpp:67-134): more than 20 neighbours, and a gradient
matrix whose determinant is below 1e-10 vol^2.  The face count is the
mesh's own (a centre cell with k faces); the singular matrix is made by
giving a cell the volume that is an eigenvalue of the geometric sum, computed
here in numpy."""
import numpy as np
import pytest

import cases
import pyoracle as po


def _cells(m):
    N = m.n_cells
    return np.vstack([np.ones(N), np.ones(N), np.zeros(N), np.zeros(N)])


@pytest.mark.parametrize("k", [20, 21, 24])
def test_neighbour_limit(k):
    m = cases.wheel_mesh(k)
    assert np.diff(m.csr_off)[0] == k
    o = po.Oracle(m, order=2)
    if k <= 20:
        o.lgi(_cells(m))
        o.flux(_cells(m))
        return
    for call in (o.lgi, o.flux):
        with pytest.raises(po.OracleError) as e:
            call(_cells(m))
        assert po.CODES[e.value.code] == "Cell has too many neighbors" and e.value.index == 0


def test_singular_matrix():
    m, i, lam = cases.singular_cell_mesh()
    c = cases.rough_cells(m, seed=2)
    c[0] = np.abs(c[0]) + 0.5
    c[1] = np.abs(c[1]) + 0.5
    with pytest.raises(po.OracleError) as e:
        po.Oracle(m, order=2).lgi(c)
    assert po.CODES[e.value.code] == "Singular matrix" and e.value.index == i
    # a volume away from the eigenvalue gives a regular matrix
    tv = m.tess_volume.copy()
    tv[i] = lam * 50
    from dataclasses import replace
    po.Oracle(replace(m, tess_volume=tv), order=2).lgi(c)
