"""Long full-size parity run of the moving mesh (not collected by pytest;
needs a GPU and a few minutes): Sedov on a 512x512 lattice in a rigid box
with Lagrangian + RoundCells points, K steps of the device-resident driver
(DeviceMovingMesh) against the same driver over the oracle's plugins
(MovingMesh over the oracle), each rebuilding its mesh every step.  Prints,
every 20 steps, whether points, dts and cells are still bit-identical.

    python tests/sedov_moving_long_parity.py [K]       (default 200)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "huji-rich_amd"), os.path.join(ROOT, "oracle")]

import cases  # noqa: E402
from rich_amd.moving import DeviceMovingMesh, MovingMesh  # noqa: E402
from test_gpu_moving import OracleBackend  # noqa: E402


def main():
    k = int(sys.argv[1]) if len(sys.argv) > 1 else 200
    m = cases.sedov_mesh(512)
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    c = cases.sedov_cells(m)
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box)
    ref = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, OracleBackend())
    t0 = time.perf_counter()
    same = True
    for step in range(1, k + 1):
        same &= dev.step() == ref.step()
        if step % 20 == 0 or step == k:
            same &= bool(np.array_equal(dev.x, ref.x) and np.array_equal(dev.y, ref.y))
            same &= bool(np.array_equal(dev.cells(), ref.cells))
            print(f"{step} steps, {time.perf_counter() - t0:.1f} s, bit-identical so far: {same}",
                  flush=True)
            if not same:
                break
    out = {"steps": step, "cells": N, "bit_identical": same,
           "max_scaled_diff_cells": float(np.max(np.abs(dev.cells() - ref.cells) /
                                                 np.maximum(np.abs(ref.cells), 1e-300)))}
    dev.close()
    print(json.dumps(out), flush=True)
    return 0 if same else 1


if __name__ == "__main__":
    sys.exit(main())
