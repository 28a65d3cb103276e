"""The C ABI from a C++ host alone (huji-rich_amd/examples/kh_driver.cpp):
one hdsim::TimeAdvance step made plugin by plugin through C-ABI calls with
host arrays in between equals the same step through rg_advance, bit for
bit, and the device loop then runs Heun steps on the same mesh.  Steps of
hdsim::TimeAdvance2Heun in the plugin adapters' call order (every state a
host array between C-ABI calls) equal the oracle's steps, bit for bit."""
import json
import os
import subprocess

import numpy as np
import pytest

import cases
import pyoracle as po

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "huji-rich_amd",
                   "bin", "kh_driver")


@pytest.mark.skipif(not os.path.exists(EXE), reason="examples not built")
@pytest.mark.parametrize("nx,ny", [(64, 32), (512, 256)])
def test_cpp_host_plugin_step_equals_device_step(nx, ny):
    out = subprocess.run([EXE, str(nx), str(ny), "20"], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["cells"] == nx * ny
    assert r["plugin_step_equals_device_step"] is True
    assert r["mcell_updates_per_s"] > 0


@pytest.mark.skipif(not os.path.exists(EXE), reason="examples not built")
@pytest.mark.parametrize("nx,ny,k", [(64, 32, 12), (256, 128, 5)])
def test_cpp_host_adapter_call_order_equals_oracle(tmp_path, nx, ny, k):
    path = str(tmp_path / "heun.bin")
    out = subprocess.run([EXE, str(nx), str(ny), "2", str(k), path], capture_output=True,
                         text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    m = cases.kh_mesh(nx, ny)
    N = m.n_cells
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == k + 12 * N
    c0 = raw[:4 * N].reshape(4, N)
    dts = raw[4 * N:4 * N + k]
    cells = raw[4 * N + k:8 * N + k].reshape(4, N)
    ext = raw[8 * N + k:].reshape(4, N)
    # the driver's initial state is the workload's (its own sin and exp)
    np.testing.assert_allclose(c0, cases.kh_cells(m), rtol=1e-14, atol=1e-17)
    o = po.Oracle(m, order=2)
    co, eo, do = o.run(c0, o.init_extensives(c0), k, integrator=2)
    np.testing.assert_array_equal(dts, do)
    np.testing.assert_array_equal(cells, co)
    np.testing.assert_array_equal(ext, eo)
