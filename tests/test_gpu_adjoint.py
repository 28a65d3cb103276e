"""GPU adjoint (rg_adjoint_forward / rg_adjoint_backward) against the CPU
forward-mode reference (oracle/tangent.cpp, dual numbers over the same
step) through the identity <J^T w, v> = <w, J v>: the adjoint of a random
final-state weight w dotted with a random initial perturbation v equals w
dotted with the tangent of v.  Both follow the branches the forward values
select, so they agree to round-off (bar: 1e-9 relative).  The forward run
under the adjoint is itself bit-identical to the oracle."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_oracle_tangent import smooth_cells

pytestmark = pytest.mark.gpu

MESHES = dict(cases.small_meshes())


def _adjoint(m, cells, nsteps, every, w, wx=None, slf=True):
    from rich_amd.gpu import RichGPU
    g = RichGPU(m, order=2, slf=slf)
    g.set_cells(cells)
    g.init_extensives()
    g.adjoint_forward(nsteps, every)
    final = g.get_cells()
    gc0 = g.adjoint_backward(w, wx)
    again = g.get_cells()
    g.close()
    return final, gc0, again


def _check(m, cells, nsteps, every, slf=True, K=3, seed=0, with_ext=False):
    rng = np.random.default_rng(seed)
    N = m.n_cells
    w = rng.standard_normal((4, N))
    wx = rng.standard_normal((4, N)) if with_ext else None
    v = rng.standard_normal((K, 4, N)) * np.abs(cells)[None]
    final, gc0, again = _adjoint(m, cells, nsteps, every, w, wx, slf)
    o = po.Oracle(m, order=2, slf=slf)
    cT, dcT, xT, dxT, _ = o.run_tangent(cells, v, nsteps)
    np.testing.assert_array_equal(final, cT)
    np.testing.assert_array_equal(again, final)
    errs = []
    for k in range(K):
        lhs = np.sum(gc0 * v[k])
        rhs = np.sum(w * dcT[k]) + (np.sum(wx * dxT[k]) if with_ext else 0.0)
        errs.append(abs(lhs - rhs) / max(abs(rhs), 1e-300))
    print(f"N={N} steps={nsteps} every={every} slf={slf} dot-product rel err max {max(errs):.2e}")
    assert max(errs) < 1e-9


@pytest.mark.parametrize("name", list(MESHES))
def test_adjoint_dot_product_small_meshes(name):
    m = MESHES[name]
    _check(m, cases.rough_cells(m, seed=31), nsteps=5, every=2)


@pytest.mark.parametrize("name", ["rigid", "periodic"])
def test_adjoint_without_limiter(name):
    m = MESHES[name]
    _check(m, smooth_cells(m), nsteps=4, every=3, slf=False)


def test_adjoint_with_extensive_seed():
    m = MESHES["periodic_flip"]
    _check(m, cases.rough_cells(m, seed=5), nsteps=3, every=1, with_ext=True)


def test_adjoint_sedov_shocked():
    """Strong blast: the shocked limiter and every HLLC wave pattern occur."""
    m = cases.sedov_mesh(32)
    _check(m, cases.sedov_cells(m), nsteps=30, every=7)


def test_adjoint_matches_gpu_finite_differences_smooth():
    """Where the step is smooth (no limiter), central differences of GPU
    forward runs agree with the adjoint."""
    from rich_amd.gpu import RichGPU
    m = MESHES["periodic"]
    cells = smooth_cells(m)
    rng = np.random.default_rng(4)
    w = rng.standard_normal((4, m.n_cells))
    _, gc0, _ = _adjoint(m, cells, 6, 4, w, slf=False)

    def loss(c):
        g = RichGPU(m, order=2, slf=False)
        g.set_cells(c)
        g.init_extensives()
        g.advance(6, integrator=2, record_dt=False)
        out = float(np.sum(w * g.get_cells()))
        g.close()
        return out

    for k in range(3):
        v = rng.standard_normal((4, m.n_cells))
        h = 1e-6
        fd = (loss(cells + h * v) - loss(cells - h * v)) / (2 * h)
        an = float(np.sum(gc0 * v))
        print("FD", fd, "adjoint", an)
        assert abs(fd - an) <= 1e-6 * abs(an)


def test_inverse_problem_gradient_against_cpu_reference():
    """The Sedov inverse problem through torch.autograd (rich_amd.inverse,
    reduced size): the directional derivatives of the ring-density loss
    w.r.t. the thermal-energy field and the blast energy equal those of the
    forward-mode CPU oracle."""
    import torch
    from rich_amd.inverse import SedovInverse
    P = SedovInverse(n=40, steps=40, every=15, r_ring=0.3, n_ring=96)
    u0 = P.u0
    E = torch.tensor(1.0, dtype=torch.float64)
    L, gu, gE = P.gradient(u0, E)
    assert L > 0
    rng = np.random.default_rng(7)
    K = 4
    dirs = [(rng.standard_normal(P.N), float(rng.standard_normal())) for _ in range(K)]
    dc = np.zeros((K, 4, P.N))
    for k, (xi, xe) in enumerate(dirs):
        dc[k, 1] = (P.gamma - 1.0) * (xi * u0.numpy() + xe * P.w.numpy())
    o = po.Oracle(P.mesh, order=2)
    c0 = P.cells(u0, E).numpy()
    cT, dcT, _, _, _ = o.run_tangent(c0, dc, P.steps)
    ring = P.ring.numpy()
    resid = cT[0, ring] - P.target.numpy()
    for k, (xi, xe) in enumerate(dirs):
        ref = float(np.sum(2.0 * resid * dcT[k, 0, ring]))
        an = float(np.sum(gu.numpy() * xi * u0.numpy())) + gE * xe
        print("adjoint", an, "forward-mode CPU", ref)
        assert abs(an - ref) <= 1e-9 * abs(ref)
    P.close()


def test_autograd_refuses_a_replaced_run():
    """A context records one forward run: after a second forward on the same
    context, the first graph's backward raises instead of returning the
    second run's gradient."""
    import torch
    from rich_amd.autograd import advance
    from rich_amd.gpu import RichGPU
    m = MESHES["periodic"]
    cells = smooth_cells(m)
    g = RichGPU(m, order=2)
    c0 = torch.from_numpy(cells).requires_grad_(True)
    out1 = advance(c0, g, 4, 2)
    out2 = advance(c0 * 1.0, g, 4, 2)
    with pytest.raises(RuntimeError, match="another forward"):
        out1.sum().backward()
    out2.sum().backward()
    assert torch.isfinite(c0.grad).all()
    g.close()


def test_adjoint_equals_finite_differences_where_no_branch_flips():
    """With the limiter on: the oracle's branch fingerprints of the runs at
    c + h v and c - h v (bit-identical to the GPU runs) tell whether the
    central difference straddles a branch switch (a jump of the limiter's
    gates or the shock switch, or a kink: an extremum, active face, wave
    pattern or CFL argmin changing hands).  Where neither run switches a
    branch in any step, the GPU central difference, taken cell by cell
    before the weighted sum, equals the adjoint to 1e-6.  Small mesh, few
    steps: few near-ties, so most directions switch nothing."""
    from rich_amd import mesh as rmesh
    from rich_amd.gpu import RichGPU
    m = rmesh.lattice(16, 12, (0.0, 1.0, 0.0, 0.75), jitter=3e-2, seed=12, bc=rmesh.PERIODIC)
    cells = smooth_cells(m)
    steps, h = 4, 1e-7
    rng = np.random.default_rng(14)
    w = rng.standard_normal((4, m.n_cells))
    _, gc0, _ = _adjoint(m, cells, steps, 2, w)
    o = po.Oracle(m, order=2)

    def fp(c):
        return o.run(c, o.init_extensives(c), steps, fingerprint=True)[3]

    def run(c):
        g = RichGPU(m, order=2)
        g.set_cells(c)
        g.init_extensives()
        g.advance(steps, integrator=2, record_dt=False)
        out = g.get_cells()
        g.close()
        return out

    base = fp(cells)
    clean = 0
    for k in range(8):
        v = rng.standard_normal((4, m.n_cells)) * np.abs(cells)
        flipped = bool(np.any(fp(cells + h * v) != base) or np.any(fp(cells - h * v) != base))
        fd = float(np.sum(w * (run(cells + h * v) - run(cells - h * v)))) / (2 * h)
        an = float(np.sum(gc0 * v))
        rel = abs(fd - an) / abs(an)
        print(f"direction {k}: branch switched {flipped}, adjoint {an:.12e} "
              f"central difference {fd:.12e} rel {rel:.2e}")
        if not flipped:
            clean += 1
            assert rel <= 1e-6
    assert clean >= 4
