"""The GPU-versus-oracle comparison itself (tests/parity.py): a NaN or an
infinity on one side only is a failure, whichever side has it."""
import math

import numpy as np
import pytest

from parity import assert_parity, compare

NAN, INF = math.nan, math.inf


@pytest.mark.parametrize("g,r", [
    ([[1, NAN, 3]], [[1, 2, 3]]),
    ([[1, 2, 3]], [[1, NAN, 3]]),
    ([[1, INF, 3]], [[1, 2, 3]]),
    ([[1, 2, 3]], [[1, -INF, 3]]),
    ([[1, INF, 3]], [[1, -INF, 3]]),
    ([[1, NAN, 3]], [[1, INF, 3]]),
])
def test_one_sided_non_finite_fails(g, r):
    err, _ = compare(g, r)
    assert err == INF
    with pytest.raises(AssertionError):
        assert_parity(g, r)


@pytest.mark.parametrize("g,r", [
    ([[1, NAN, 3]], [[1, NAN, 3]]),
    ([[1, INF, 3]], [[1, INF, 3]]),
    ([[-INF, 2.0]], [[-INF, 2.0]]),
])
def test_matching_non_finite_passes(g, r):
    err, _ = compare(g, r)
    assert err == 0.0
    assert_parity(g, r)


def test_scaled_error_and_bit_fraction():
    r = np.array([[1.0, 2.0, 1e-9, 4.0]])
    g = r.copy()
    g[0, 1] = 2.0 * (1 + 1e-13)
    g[0, 2] = 2e-9  # measured against 1e-3 of the row's largest value
    err, same = compare(g, r)
    assert err == pytest.approx(max(1e-13, 1e-9 / 4e-3), rel=1e-3)
    assert same == 0.5
    with pytest.raises(AssertionError):
        assert_parity(g, r)
    assert_parity(g, r, rtol=1e-6)


def test_empty():
    assert compare(np.zeros((4, 0)), np.zeros((4, 0))) == (0.0, 1.0)
