"""Long full-size parity run (not collected by pytest; needs a GPU and a few
minutes): Kelvin-Helmholtz on the 2048x1024 periodic jittered lattice,
advanced K second-order Heun steps on the GPU and by the oracle's OpenMP
form from the same state; prints whether the time steps, cells and
extensives are bit-identical, and the largest scaled difference.

    python tests/kh_long_parity.py [K]        (default 300)
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
import pyoracle as po  # noqa: E402
from rich_amd.gpu import RichGPU  # noqa: E402


def main():
    k = int(sys.argv[1]) if len(sys.argv) > 1 else 300
    m = cases.kh_mesh(2048, 1024)
    cells = cases.kh_cells(m)
    t0 = time.perf_counter()
    g = RichGPU(m, order=2)
    g.set_cells(cells)
    g.init_extensives()
    dts = g.advance(k, integrator=2)
    cg, xg = g.get_cells(), g.get_extensives()
    g.close()
    t1 = time.perf_counter()
    print(f"GPU: {k} steps in {t1 - t0:.1f} s", flush=True)
    # the oracle in segments of 100 steps (each restarts from the last
    # cells and extensives, as the step itself does), printing progress
    o = po.Oracle(m, order=2, nthreads=min(16, os.cpu_count() or 1))
    co, xo, parts = cells, o.init_extensives(cells), []
    for s0 in range(0, k, 100):
        co, xo, d = o.run(co, xo, min(100, k - s0), integrator=2)
        parts.append(d)
        print(f"oracle: {s0 + len(d)} steps, {time.perf_counter() - t1:.1f} s", flush=True)
    do = np.concatenate(parts)

    def scaled(a, b):
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))

    out = {"steps": k, "cells": m.n_cells,
           "dt_identical": bool(np.array_equal(dts, do)),
           "cells_identical": bool(np.array_equal(cg, co)),
           "extensives_identical": bool(np.array_equal(xg, xo)),
           "max_scaled_diff_cells": scaled(cg, co), "sim_time": float(np.sum(do))}
    print(json.dumps(out), flush=True)
    return 0 if out["dt_identical"] and out["cells_identical"] and out["extensives_identical"] else 1


if __name__ == "__main__":
    sys.exit(main())
