"""Long full-size parity run of the batched ensemble (not collected by
pytest; needs a GPU and a few minutes): the 256 Sedov runs of 66^2..255^2
cells (8.4M cells), K steps in one graph launch per step on the GPU, against
each instance run alone by the oracle's OpenMP form.  Prints progress every
32 instances and whether every instance's dts and cells are bit-identical.

    python tests/ensemble_long_parity.py [K]       (default 100)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "huji-rich_amd"), os.path.join(ROOT, "oracle")]

import pyoracle as po  # noqa: E402
from rich_amd import ensemble  # noqa: E402
from rich_amd.gpu import RichGPU  # noqa: E402


def main():
    k = int(sys.argv[1]) if len(sys.argv) > 1 else 100
    ms, cells, _ = ensemble.sedov_ensemble(K=256)
    u = ensemble.union(ms)
    t0 = time.perf_counter()
    g = RichGPU(u.mesh, order=2)
    g.set_cells(np.concatenate(cells, axis=1))
    g.init_extensives()
    g.set_instances(u.cell_off)
    dts = g.advance_ensemble(k)
    out = u.split(g.get_cells())
    g.close()
    print(f"GPU: {len(ms)} instances, {k} steps, {time.perf_counter() - t0:.1f} s", flush=True)
    bad = []
    for j, m in enumerate(ms):
        o = po.Oracle(m, order=2, nthreads=min(16, os.cpu_count() or 1))
        co, _, do = o.run(cells[j], o.init_extensives(cells[j]), k, integrator=2)
        if not (np.array_equal(dts[:, j], do) and np.array_equal(out[j], co)):
            bad.append(j)
        if (j + 1) % 32 == 0:
            print(f"oracle: {j + 1} instances, {time.perf_counter() - t0:.1f} s, "
                  f"mismatched so far: {len(bad)}", flush=True)
    res = {"instances": len(ms), "steps": k, "cells": int(u.cell_off[-1]),
           "bit_identical_instances": len(ms) - len(bad), "mismatched": bad[:10]}
    print(json.dumps(res), flush=True)
    return 0 if not bad else 1


if __name__ == "__main__":
    sys.exit(main())
