"""The measured workloads at their full sizes, checked through what does not
depend on size: the GPU step equals the oracle on the same mesh and state
(bit for bit, a few steps), totals are conserved, the mixed-precision step
stays within its error bound of FP64, the ensemble's instances equal their
runs alone, and the adjoint satisfies the dot-product identity against the
forward-mode oracle.

Workloads: Kelvin-Helmholtz on a periodic 2048x1024 jittered lattice
(2,097,152 cells); Sedov on a moving 512x512 mesh (262,144 cells); an
ensemble of 256 Sedov runs of 64^2..256^2 cells (about 8.4M cells); the
Sedov adjoint on 256x256 (65,536 cells)."""
import numpy as np
import pytest

import cases
import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kh():
    m = cases.kh_mesh(2048, 1024)
    return m, cases.kh_cells(m)


def test_kh_2048x1024_equals_oracle(kh):
    from rich_amd.gpu import RichGPU
    m, cells = kh
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    g.init_extensives()
    dts = g.advance(2, integrator=2)
    o = po.Oracle(m, order=2, nthreads=16)
    co, eo, do = o.run(cells, o.init_extensives(cells), 2, integrator=2)
    np.testing.assert_array_equal(dts, do)
    np.testing.assert_array_equal(g.get_extensives(), eo)
    np.testing.assert_array_equal(g.get_cells(), co)
    g.close()


def test_kh_2048x1024_conservation_and_mixed_precision(kh):
    """Periodic box: mass, energy and momentum totals are conserved to
    round-off over 50 steps in FP64 and in mixed precision; the mixed run's
    density stays within 1e-5 relative L1 of FP64."""
    from rich_amd.gpu import RG_MIXED, RichGPU
    m, cells = kh
    out = {}
    for name, prec in (("fp64", None), ("mixed", RG_MIXED)):
        g = RichGPU(m, order=2, **({} if prec is None else {"precision": prec}))
        g.set_cells(cells)
        g.init_extensives()
        x0 = g.get_extensives()
        g.advance(50, integrator=2, record_dt=False)
        x1 = g.get_extensives()
        # momentum against the momentum scale sum sqrt(2 m E): the y momentum
        # itself is only the small seed perturbation
        pscale = np.sqrt(2 * x0[0] * x0[1]).sum()
        for k in range(4):
            scale = np.abs(x0[k]).sum() if k < 2 else pscale
            assert abs(x1[k].sum() - x0[k].sum()) <= 1e-13 * scale, (name, k)
        out[name] = g.get_cells()
        g.close()
    d64, dm = out["fp64"][0], out["mixed"][0]
    l1 = np.abs(dm - d64).sum() / np.abs(d64).sum()
    print(f"mixed vs FP64 density L1 after 50 steps: {l1:.2e}")
    assert l1 <= 1e-5


def test_moving_sedov_512x512_device_equals_oracle():
    """Two moving-mesh steps at full size: the device-resident step and the
    same driver over the oracle give the same points, dts and state."""
    from rich_amd.moving import DeviceMovingMesh
    from oracle_drivers import oracle_moving_mesh
    m = cases.sedov_mesh(512)
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    c = cases.sedov_cells(m)
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box)
    ref = oracle_moving_mesh(m.pt_x[:N], m.pt_y[:N], c, box)
    x0 = dev.extensives()
    for step in range(2):
        assert dev.step() == ref.step(), step
        np.testing.assert_array_equal(dev.x, ref.x)
        np.testing.assert_array_equal(dev.y, ref.y)
        np.testing.assert_array_equal(dev.cells(), ref.cells)
    x1 = dev.extensives()
    assert abs(x1[0].sum() / x0[0].sum() - 1) < 1e-14
    assert abs(x1[1].sum() / x0[1].sum() - 1) < 1e-14
    dev.close()


def test_ensemble_256_sedov_instances_equal_oracle():
    """The full ensemble, 3 steps in one graph launch each; sampled
    instances (the smallest, the largest and two others) equal the oracle
    run of that instance alone, dts included."""
    from rich_amd import ensemble
    from rich_amd.gpu import RichGPU
    ms, cells, prm = ensemble.sedov_ensemble(K=256)
    u = ensemble.union(ms)
    assert u.cell_off[-1] > 8_000_000
    g = RichGPU(u.mesh, order=2)
    g.set_cells(np.concatenate(cells, axis=1))
    g.init_extensives()
    g.set_instances(u.cell_off)
    dts = g.advance_ensemble(3)
    out = u.split(g.get_cells())
    g.close()
    n = np.array([mm.n_cells for mm in ms])
    for k in sorted({int(np.argmin(n)), int(np.argmax(n)), 17, 200}):
        o = po.Oracle(ms[k], order=2, nthreads=16)
        co, eo, do = o.run(cells[k], o.init_extensives(cells[k]), 3, integrator=2)
        np.testing.assert_array_equal(dts[:, k], do)
        np.testing.assert_array_equal(out[k], co)


def test_adjoint_sedov_256x256_dot_product():
    """Adjoint of 12 steps (checkpoint every 5) on the full 256x256 mesh
    against the forward-mode oracle: <J^T w, v> = <w, J v>."""
    from rich_amd.gpu import RichGPU
    m = cases.sedov_mesh(256)
    cells = cases.sedov_cells(m)
    rng = np.random.default_rng(21)
    w = rng.standard_normal((4, m.n_cells))
    v = rng.standard_normal((2, 4, m.n_cells)) * np.abs(cells)[None]
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    g.init_extensives()
    g.adjoint_forward(12, 5)
    final = g.get_cells()
    gc0 = g.adjoint_backward(w)
    g.close()
    cT, dcT, _, _, _ = po.Oracle(m, order=2, nthreads=16).run_tangent(cells, v, 12)
    np.testing.assert_array_equal(final, cT)
    for k in range(2):
        lhs, rhs = np.sum(gc0 * v[k]), np.sum(w * dcT[k])
        print(f"direction {k}: adjoint {lhs:.15e} forward-mode {rhs:.15e}")
        assert abs(lhs - rhs) <= 1e-9 * abs(rhs)


def test_adjoint_256x256_matches_finite_differences_without_limiter():
    """Where the step is differentiable (limiter off, smooth state), central
    differences of full-size GPU runs agree with the adjoint to 1e-6
    relative; 40 steps, checkpoint every 20.  The difference is taken cell by
    cell before the weighted sum, so that the sum does not cancel two large
    nearly equal losses.  Even so the step is only piecewise smooth: the CFL
    dt is a minimum over 65k cells (near ties switch its argmin) and the HLLC
    wave pattern switches where a velocity crosses a wave speed, so one step
    size may straddle a kink.  Each direction is differenced with h = 1e-6
    and 1e-7 and the closer of the two must agree."""
    from rich_amd.gpu import RichGPU
    from test_oracle_tangent import smooth_cells
    m = cases.sedov_mesh(256)
    cells = smooth_cells(m)
    rng = np.random.default_rng(33)
    w = rng.standard_normal((4, m.n_cells))
    g = RichGPU(m, order=2, slf=False)
    g.set_cells(cells)
    g.init_extensives()
    g.adjoint_forward(40, 20)
    gc0 = g.adjoint_backward(w)
    g.close()

    def run(c):
        gg = RichGPU(m, order=2, slf=False)
        gg.set_cells(c)
        gg.init_extensives()
        gg.advance(40, integrator=2, record_dt=False)
        out = gg.get_cells()
        gg.close()
        return out

    for k in range(3):
        v = rng.standard_normal((4, m.n_cells)) * np.abs(cells)
        an = float(np.sum(gc0 * v))
        errs = []
        for h in (1e-6, 1e-7):
            fd = float(np.sum(w * (run(cells + h * v) - run(cells - h * v)))) / (2 * h)
            errs.append(abs(fd - an) / abs(an))
            print(f"direction {k} h {h:.0e}: adjoint {an:.12e} central difference {fd:.12e} "
                  f"rel {errs[-1]:.2e}")
        assert min(errs) <= 1e-6
