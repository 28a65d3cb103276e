"""The C ABI from a C++ host alone (huji-rich_amd/examples/kh_driver.cpp):
one hdsim::TimeAdvance step made plugin by plugin through C-ABI calls with
host arrays in between equals the same step through rg_advance, bit for
bit, and the device loop then runs Heun steps on the same mesh."""
import json
import os
import subprocess

import pytest

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
