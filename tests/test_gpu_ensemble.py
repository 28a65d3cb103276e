"""Batched ensemble on the GPU (rg_set_instances, one hipGraph launch of
persistent kernels per step): every instance's state and time steps are
bit-identical to running that instance alone (itself bit-identical to the
oracle), for every launch shape the autotuner tries."""
import numpy as np
import pytest

import pyoracle as po
from rich_amd import ensemble

pytestmark = pytest.mark.gpu

NSTEPS = 12


def _ensemble(K=5, seed=3):
    ms, cells, prm = ensemble.sedov_ensemble(K=K, n_min=8, n_max=20, seed=seed)
    return ms, cells, prm, ensemble.union(ms)


def _alone(m, c, nsteps):
    from rich_amd.gpu import RichGPU
    g = RichGPU(m, order=2)
    g.set_cells(c)
    g.init_extensives()
    dts = g.advance(nsteps, integrator=2)
    out = g.get_cells()
    g.close()
    return out, dts


def _batched(u, cells, nsteps, shape=None, tune=False):
    from rich_amd.gpu import RichGPU
    g = RichGPU(u.mesh, order=2)
    g.set_cells(np.concatenate(cells, axis=1))
    g.init_extensives()
    g.set_instances(u.cell_off)
    table = None
    if shape:
        g.ensemble_config(*shape)
    if tune:
        table = g.ensemble_autotune(trial_steps=2)
    dts = g.advance_ensemble(nsteps)
    out = g.get_cells()
    g.close()
    return u.split(out), dts, table


def test_ensemble_equals_instances_alone():
    ms, cells, prm, u = _ensemble()
    got, dts, _ = _batched(u, cells, NSTEPS)
    assert dts.shape == (NSTEPS, len(ms))
    for k, (m, c) in enumerate(zip(ms, cells)):
        ref, rdt = _alone(m, c, NSTEPS)
        np.testing.assert_array_equal(got[k], ref)
        np.testing.assert_array_equal(dts[:, k], rdt)
    # instances with different energies have different time steps
    assert len(set(np.round(dts[0], 15))) > 1
    # one instance against the oracle
    o = po.Oracle(ms[0], order=2)
    c0, _, d0 = o.run(cells[0], o.init_extensives(cells[0]), NSTEPS)
    np.testing.assert_array_equal(got[0], c0)
    np.testing.assert_array_equal(dts[:, 0], d0)


@pytest.mark.parametrize("shape", [(64, 1), (128, 3), (256, 16), (128, 0)])
def test_launch_shape_does_not_change_results(shape):
    ms, cells, prm, u = _ensemble(K=4, seed=5)
    base, bdts, _ = _batched(u, cells, 6)
    got, dts, _ = _batched(u, cells, 6, shape=shape)
    for a, b in zip(base, got):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(bdts, dts)


def test_autotune_restores_state_and_keeps_results():
    ms, cells, prm, u = _ensemble(K=4, seed=6)
    base, bdts, _ = _batched(u, cells, 6)
    got, dts, table = _batched(u, cells, 6, tune=True)
    assert table and all(ms_ > 0 for _, _, ms_ in table)
    print("autotune table", table)
    for a, b in zip(base, got):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(bdts, dts)
