"""Full-length parity of the measured workloads (GPU, several minutes of
oracle time on the host cores), collected after the other GPU tests:

* Kelvin-Helmholtz 2048x1024, 1000 Heun steps: every dt, cell and
  extensive bit-identical to the oracle's OpenMP form, compared every 100
  steps;
* moving Sedov 512x512, 200 steps with a mesh rebuild every step: the
  device-resident run against the same run on the oracle (plugins and
  point motion), points, dts and cells compared every 20 steps;
* the batched ensemble of 256 Sedov runs (8.4M cells), 100 steps in one
  graph launch per step: every instance's dts and cells equal its run
  alone through the oracle;
* the Sedov adjoint as specified (256x256, 400 steps, checkpoint every 20,
  ring-density loss at r = 0.3 against the E* = 1.3 target): dL/du and dL/dE
  against the forward-mode oracle along seeded directions.

Each test appends progress lines to gpurun_out/long_parity_progress.txt
while the oracle runs."""
import os
import time

import numpy as np
import pytest

import cases
import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREADS = min(16, os.cpu_count() or 1)


def progress(msg):
    d = os.path.join(ROOT, "gpurun_out")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "long_parity_progress.txt"), "a") as f:
        f.write(f"{time.strftime('%H:%M:%S')} {msg}\n")


@pytest.mark.timeout(1200)
def test_kh_2048x1024_1000_steps():
    from rich_amd.gpu import RichGPU
    m = cases.kh_mesh(2048, 1024)
    cells = cases.kh_cells(m)
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    g.init_extensives()
    o = po.Oracle(m, order=2, nthreads=THREADS)
    co, xo = cells, o.init_extensives(cells)
    for s0 in range(0, 1000, 100):
        dts = g.advance(100, integrator=2)
        co, xo, do = o.run(co, xo, 100, integrator=2)
        np.testing.assert_array_equal(dts, do, err_msg=f"dt, steps {s0}..{s0 + 99}")
        np.testing.assert_array_equal(g.get_cells(), co, err_msg=f"cells after {s0 + 100}")
        np.testing.assert_array_equal(g.get_extensives(), xo, err_msg=f"ext after {s0 + 100}")
        progress(f"kh 2048x1024: {s0 + 100} steps bit-identical")
    g.close()


@pytest.mark.timeout(1200)
def test_moving_sedov_512_200_steps():
    from oracle_drivers import oracle_moving_mesh
    from rich_amd.moving import DeviceMovingMesh
    m = cases.sedov_mesh(512)
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    c = cases.sedov_cells(m)
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box)
    ref = oracle_moving_mesh(m.pt_x[:N], m.pt_y[:N], c, box)
    for step in range(1, 201):
        assert dev.step() == ref.step(), step
        if step % 20 == 0:
            np.testing.assert_array_equal(dev.x, ref.x)
            np.testing.assert_array_equal(dev.y, ref.y)
            np.testing.assert_array_equal(dev.cells(), ref.cells)
            np.testing.assert_array_equal(dev.extensives(), ref.ext)
            progress(f"moving sedov 512: {step} steps bit-identical")
    dev.close()


@pytest.mark.timeout(1200)
def test_ensemble_256_100_steps():
    from rich_amd import ensemble
    from rich_amd.gpu import RichGPU
    ms, cells, _ = ensemble.sedov_ensemble(K=256)
    u = ensemble.union(ms)
    g = RichGPU(u.mesh, order=2)
    g.set_cells(np.concatenate(cells, axis=1))
    g.init_extensives()
    g.set_instances(u.cell_off)
    dts = g.advance_ensemble(100)
    out = u.split(g.get_cells())
    g.close()
    bad = []
    for j, mj in enumerate(ms):
        o = po.Oracle(mj, order=2, nthreads=THREADS)
        co, _, do = o.run(cells[j], o.init_extensives(cells[j]), 100, integrator=2)
        if not (np.array_equal(dts[:, j], do) and np.array_equal(out[j], co)):
            bad.append(j)
        if (j + 1) % 32 == 0:
            progress(f"ensemble: {j + 1} instances checked, {len(bad)} mismatched")
    assert not bad, bad[:10]


@pytest.mark.timeout(1200)
def test_sedov_adjoint_as_specified():
    import torch
    from rich_amd.inverse import SedovInverse
    P = SedovInverse(n=256, steps=400, every=20, r_ring=0.3, n_ring=1024, E_target=1.3)
    u0 = P.u0
    E = torch.tensor(1.0, dtype=torch.float64)
    L, gu, gE = P.gradient(u0, E)
    progress(f"adjoint C5: loss {L:.3e}, dL/dE {gE:.3e}")
    rng = np.random.default_rng(5)
    dirs = [(rng.standard_normal(P.N), float(rng.standard_normal())) for _ in range(2)]
    dirs.append((np.zeros(P.N), 1.0))  # dL/dE alone
    dc = np.zeros((len(dirs), 4, P.N))
    for k, (xi, xe) in enumerate(dirs):
        dc[k, 1] = (P.gamma - 1.0) * (xi * u0.numpy() + xe * P.w.numpy())
    o = po.Oracle(P.mesh, order=2, nthreads=THREADS)
    c0 = P.cells(u0, E).numpy()
    cT, dcT, _, _, _ = o.run_tangent(c0, dc, P.steps)
    ring = P.ring.numpy()
    resid = cT[0, ring] - P.target.numpy()
    for k, (xi, xe) in enumerate(dirs):
        ref = float(np.sum(2.0 * resid * dcT[k, 0, ring]))
        an = float(np.sum(gu.numpy() * xi * u0.numpy())) + gE * xe
        progress(f"adjoint C5 direction {k}: adjoint {an:.15e} forward-mode {ref:.15e}")
        assert ref != 0.0
        assert abs(an - ref) <= 1e-9 * abs(ref)
    P.close()
