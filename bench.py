#!/usr/bin/env python3
"""Throughput of the device-resident RICH finite-volume step on one MI355X.

Workload: 2D Kelvin-Helmholtz, 2048 x 1024 jittered lattice (2,097,152
cells) in a periodic [0,2]x[0,1] box, second order (LinearGaussImproved +
slope limiter, HLLC, Heun), static mesh, FP64.  One step = one
hdsim::TimeAdvance2Heun pass (two flux stages) over all cells; inputs are
resident in HBM before the timed region.

Prints ONE JSON line (rank 0).  `value` is Mcell-updates/s (cells x steps /
s).  `roofline` is one flux stage (k_slopes, k_face_update -- the faces
and the cell update fused per tile of 256 cells -- and k_reduce_min; a Heun
step has two): the algorithmic bytes of a stage, 180 N + 48 F (each datum
of the step read or written once), over the stage's summed HIP-event kernel
time; `traffic` is the HBM bytes the counters measured for the same
kernels.  `kernels` gives each kernel's
time and its own model bytes (every array it touches, once), and
`stage_roofline` the same algorithmic bytes over the timed wall clock.
`cpu_baseline` times the CPU restatement of the reference step (oracle/,
OpenMP) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "huji-rich_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


STAGE_KERNELS = ("k_slopes", "k_face_update", "k_reduce_min")
TILE = 256  # cells per tile of k_face_update


def tile_entries(mesh):
    """Faces k_face_update computes per stage: each face once per tile that
    holds one of its real sides."""
    import numpy as np
    N = mesh.n_cells
    a, b = np.asarray(mesh.face_first), np.asarray(mesh.face_second)
    ra, rb = a < N, b < N
    same = ra & rb & (a // TILE == b // TILE)
    return int(np.count_nonzero(ra) + np.count_nonzero(rb) - np.count_nonzero(same))


def kernel_alg_bytes(name, N, F, G, nnz, moving=False, entries=0):
    """Model bytes of one launch: every array the kernel's own inputs and
    outputs touch, each datum once (FP64 values, int32 indices).  Summed over
    a stage's kernels this counts the intermediates (slopes, fluxes) and the
    raw geometry several times; the stage's algorithmic bytes are
    180 N + 48 F."""
    if name == "k_slopes":
        # cells 32N, ghost sources 4G, CSR 4(N+1)+8nnz, CM+points 32(N+G),
        # tess volume 8N, face vertices 32F; slopes out 64N
        return 32 * N + 4 * G + 4 * (N + 1) + 8 * nnz + 32 * (N + G) + 8 * N + 32 * F + 64 * N
    if name == "k_flux2":
        # face pair 8F, vertices 32F, cells 32N, slopes 64N, CM+points
        # 32(N+G), ghost ids 8G, face velocity 16F if moving; flux out 32F
        return 8 * F + 32 * F + 32 * N + 64 * N + 32 * (N + G) + 8 * G + (16 * F if moving else 0) + 32 * F
    if name == "k_update":
        # extensives in 32N, Heun average 32N/2 (one of two stages), flux 32F,
        # area 8F, CSR 4(N+1)+4nnz, volumes 16N; extensives + cells out 64N
        return 32 * N + 16 * N + 32 * F + 8 * F + 4 * (N + 1) + 4 * nnz + 16 * N + 64 * N
    if name == "k_face_update":
        # per tile entry: face frame 48 and sides 8; cells 32N, slopes 64N,
        # CM 16(N+G), ghost ids 8G; extensives in 32N, Heun average 32N/2,
        # per-slot entry words 2nnz, volumes 16N; extensives + cells out 64N
        return (56 * entries + 32 * N + 64 * N + 16 * (N + G) + 8 * G
                + 32 * N + 16 * N + 2 * nnz + 16 * N + 64 * N)
    if name == "k_flux1":
        return 8 * F + 32 * F + 32 * N + 32 * (N + G) + 32 * F
    if name == "k_cfl":
        return 32 * N + 8 * N + 4 * (N + 1)
    return 0


def measured_traffic(kernel, nx, ny):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, calibrated); None if
    there is no measurement for it."""
    path = os.path.join(ROOT, "profiles", f"kh{nx}x{ny}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    i = text.find("```json")
    if i < 0:
        return None
    block = text[i + 7:text.index("```", i + 7)]
    return json.loads(block).get("kernels", {}).get(kernel)


def calibrate_copy(mib):
    """One hipMemcpy device-to-device of `mib` MiB (plus one warm-up copy of
    1 MiB), through the HIP runtime: its copy kernel reads and writes exactly
    mib MiB, which calibrates FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(1 << 20), 3)
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # 3 = hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": n}))


def cpu_baseline(mesh, cells, steps, threads):
    import pyoracle as po
    o = po.Oracle(mesh, order=2, nthreads=threads)
    ext = o.init_extensives(cells)
    t0 = time.perf_counter()
    c, e, dts = o.run(cells, ext, steps, integrator=2)
    dt = time.perf_counter() - t0
    return c, e, dts, dt


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def moving_leg(args):
    """Sedov on a moving mesh: Lagrangian points with RoundCells, the Voronoi
    mesh rebuilt on the host every step.  The state stays on the GPU
    (DeviceMovingMesh); the same steps driven through the plugin-level ABI
    (MovingMesh over GpuBackend, every state copied each way) are timed
    beside it."""
    import numpy as np
    import cases
    from rich_amd.moving import DeviceMovingMesh, GpuBackend, MovingMesh

    n = args.moving_n
    steps = args.moving_steps
    log(f"moving mesh: Sedov {n}x{n}, {steps} steps")
    m = cases.sedov_mesh(n)
    c = cases.sedov_cells(m)
    N = m.n_cells
    box = (-1.0, 1.0, -1.0, 1.0)
    t0 = time.perf_counter()
    dev = DeviceMovingMesh(m.pt_x[:N], m.pt_y[:N], c, box)
    setup = time.perf_counter() - t0
    x0 = dev.extensives()
    t0 = time.perf_counter()
    for k in range(steps):
        dev.step()
        if (k + 1) % 20 == 0:
            log(f"moving mesh: step {k + 1}, {time.perf_counter() - t0:.1f} s")
    el = time.perf_counter() - t0
    x1 = dev.extensives()
    T = dev.timing
    out = {"workload": f"Sedov {n}x{n} jittered lattice in [-1,1]^2, rigid walls, Lagrangian + "
                       "RoundCells points, Voronoi rebuilt on the host every step, "
                       "LinearGaussImproved + HLLC, Heun, FP64",
           "cells": N, "steps": steps, "sim_time": dev.time,
           "s_per_step": el / steps, "value": N * steps / el / 1e6, "unit": "Mcell-updates/s",
           "host_mesh_s_per_step": T["mesh_s"] / steps,
           "begin_s_per_step": T["begin_s"] / steps,
           "replace_mesh_s_per_step": T["replace_s"] / steps,
           "end_s_per_step": T["end_s"] / steps,
           "mesh_builds_repairs": list(dev.mesher.counts()),
           "hilbert_resort": {"every": dev.resort_every, "reorders": dev.reorders},
           "mass_drift": float(x1[0].sum() / x0[0].sum() - 1.0),
           "energy_drift": float(x1[1].sum() / x0[1].sum() - 1.0),
           "setup_s": setup,
           "note": "begin: point velocities, dt and stage 1 on the device, point velocities "
                   "copied back; replace: the rebuilt mesh's tables built and uploaded; end: "
                   "stage 2 on the device"}
    dev.close()
    # the plugin-level flow over the first steps only (it is the slower one)
    psteps = min(steps, 10)
    run = MovingMesh(m.pt_x[:N], m.pt_y[:N], c, box, GpuBackend())
    run.timing = {k: 0.0 for k in run.timing}
    t0 = time.perf_counter()
    for k in range(psteps):
        run.step()
    el = time.perf_counter() - t0
    T = run.timing
    out["plugin_abi"] = {
        "steps": psteps, "s_per_step": el / psteps, "value": N * psteps / el / 1e6,
        "host_mesh_s_per_step": T["mesh_s"] / psteps,
        "host_motion_s_per_step": T["motion_s"] / psteps,
        "gpu_plugins_s_per_step": T["plugins_s"] / psteps,
        "note": "plugin-level C ABI: cells, extensives and fluxes copied each way twice per "
                "step, one context and mesh upload per step"}
    run.b.close()
    return out


def ensemble_leg(args):
    """256 independent Sedov runs of uneven size advanced together (one
    hipGraph launch per step, per-instance dt, the slope kernel's launch
    shape from the autotuner, faces and cell update fused per tile), against
    the same instances advanced one at a time."""
    import numpy as np
    from rich_amd import ensemble
    from rich_amd.gpu import RichGPU

    log(f"ensemble: building {args.ensemble_k} instances")
    t0 = time.perf_counter()
    ms, cells, prm = ensemble.sedov_ensemble(K=args.ensemble_k)
    u = ensemble.union(ms)
    g = RichGPU(u.mesh, order=2, cfl=0.3)
    g.set_cells(np.concatenate(cells, axis=1))
    g.init_extensives()
    g.set_instances(u.cell_off)
    setup = time.perf_counter() - t0
    log(f"ensemble: {int(u.cell_off[-1])} cells set up in {setup:.1f} s; autotuning")
    table = g.ensemble_autotune(trial_steps=3)
    best = min(table, key=lambda r: r[2])
    steps = args.ensemble_steps
    g.advance_ensemble(2, record_dt=False)  # warm-up
    g.set_cells(np.concatenate(cells, axis=1))
    g.init_extensives()
    t = time.perf_counter()
    dts = g.advance_ensemble(steps)
    el = time.perf_counter() - t
    final = u.split(g.get_cells())
    fused, entries, _ = g.tiles()
    g.close()
    N = int(u.cell_off[-1])
    # a sample advanced alone: the same bits, and the sequential rate
    sample = list(range(0, len(ms), max(1, len(ms) // 8)))[:8]
    same, t_alone, n_alone = 0, 0.0, 0
    for k in sample:
        a = RichGPU(ms[k], order=2, cfl=0.3)
        a.set_cells(cells[k])
        a.init_extensives()
        a.advance(2, integrator=2, record_dt=False)
        a.set_cells(cells[k])
        a.init_extensives()
        t = time.perf_counter()
        adt = a.advance(steps, integrator=2)
        t_alone += time.perf_counter() - t
        n_alone += ms[k].n_cells
        same += int(np.array_equal(a.get_cells(), final[k]) and np.array_equal(adt, dts[:, k]))
        a.close()
    value = N * steps / el / 1e6
    seq_rate = n_alone * steps / t_alone / 1e6
    return {"workload": f"{len(ms)} independent Sedov runs, square jittered lattices "
                        f"{int(prm['n'].min())}^2..{int(prm['n'].max())}^2 cells in [-1,1]^2, "
                        "rigid walls, E log-uniform [0.1,10], deposit radius 3 cell spacings, "
                        "second order, Heun, static mesh, FP64, per-instance CFL dt",
            "instances": len(ms), "cells_total": N, "faces_total": int(u.mesh.n_faces),
            "steps": steps, "value": value, "unit": "Mcell-updates/s",
            "ms_per_step": 1e3 * el / steps,
            "launch": {"block": best[0], "blocks_per_cu": best[1], "graph_launches_per_step": 1,
                       "kernels_per_step": 6 if fused else 8,
                       "fused_stages": fused, "tile_entries": entries},
            "autotune_ms_per_step": [[b, p, round(t_, 4)] for b, p, t_ in table],
            "one_at_a_time": {"instances": len(sample), "value": seq_rate,
                              "batched_speedup": value / seq_rate},
            "bit_identical_to_alone": f"{same}/{len(sample)}",
            "dt_range": [float(dts[0].min()), float(dts[0].max())],
            "setup_s": setup}


def adjoint_leg(args):
    """Sedov inverse problem (static rigid box): timings of the checkpointed
    forward and of the adjoint, and the adjoint's directional derivatives
    against central differences (h = 1e-6, directions relative to the
    parameters; each difference taken cell by cell before the loss) and
    optionally against the forward-mode CPU oracle.  For the first
    --adjoint-fp-dirs directions and the blast energy alone, the oracle's
    branch fingerprints of the runs at -h and +h (bit-identical to the GPU
    runs) record which steps switch a branch between them: a jump (limiter
    gate or guard, shock switch) or a kink (extremum neighbour, active face,
    HLLC wave pattern, CFL argmin)."""
    import numpy as np
    import torch
    from rich_amd.inverse import SedovInverse

    n, steps, every = args.adjoint_n, args.adjoint_steps, args.adjoint_every
    log(f"adjoint: Sedov {n}x{n}, {steps} steps")
    P = SedovInverse(n=n, steps=steps, every=every)
    s = P.solver
    u0 = P.u0
    E = torch.tensor(1.0, dtype=torch.float64)
    c0 = P.cells(u0, E).numpy().copy()
    w = np.random.default_rng(1).standard_normal((4, P.N))
    plain, fck, bwd = [], [], []
    for _ in range(3):
        s.set_cells(c0)
        s.init_extensives()
        t = time.perf_counter()
        s.advance(steps, integrator=2, record_dt=False)
        plain.append(time.perf_counter() - t)
        s.set_cells(c0)
        s.init_extensives()
        t = time.perf_counter()
        s.adjoint_forward(steps, every, record_dt=False)
        fck.append(time.perf_counter() - t)
        t = time.perf_counter()
        s.adjoint_backward(w)
        bwd.append(time.perf_counter() - t)
    tf, tc, tb = (float(np.median(x)) for x in (plain, fck, bwd))
    out = {"workload": f"Sedov {n}x{n} jittered lattice, rigid box [-1,1]^2, static mesh, "
                       f"{steps} Heun steps, checkpoint every {every}, FP64",
           "cells": P.N, "forward_ms": 1e3 * tf, "forward_checkpointed_ms": 1e3 * tc,
           "backward_ms": 1e3 * tb, "backward_over_forward": tb / tf,
           "forward_and_backward_over_forward": (tc + tb) / tf}
    rng = np.random.default_rng(0)
    dirs = [(torch.from_numpy(rng.standard_normal(P.N)), float(rng.standard_normal()))
            for _ in range(args.adjoint_dirs)]
    e_only = (torch.zeros(P.N, dtype=torch.float64), 1.0)
    h = 1e-6

    def final_cells(xi, xe, sign):
        with torch.no_grad():
            return P.forward(u0 + sign * h * xi * u0, E + sign * h * xe * E)[1].numpy()

    # each direction's runs at -h and +h serve both rings
    pairs = [(final_cells(xi, xe, -1.0), final_cells(xi, xe, 1.0)) for xi, xe in dirs + [e_only]]
    log("adjoint: finite-difference runs done")
    # branch fingerprints of the oracle runs at -h, 0, +h
    fp = None
    kfp = min(args.adjoint_fp_dirs, len(dirs))
    if kfp > 0:
        import pyoracle as po
        threads = min(16, os.cpu_count() or 1)
        o = po.Oracle(P.mesh, order=2, nthreads=threads)

        def fprint(xi, xe, sign):
            c = P.cells(u0 + sign * h * xi * u0, E + sign * h * xe * E).numpy()
            return o.run(c, o.init_extensives(c), steps, fingerprint=True)[3]

        t = time.perf_counter()
        base = fprint(e_only[0], 0.0, 0.0)
        fp = []
        for xi, xe in dirs[:kfp] + [e_only]:
            diff = (fprint(xi, xe, -1.0) != base) | (fprint(xi, xe, 1.0) != base)
            fp.append({"steps_with_jumps": int(diff[:, 6:8].any(axis=1).sum()),
                       "steps_with_kinks": int(diff[:, 2:6].any(axis=1).sum())})
        log(f"adjoint: branch fingerprints of {kfp} directions and E done")
        fp_meta = {"directions": kfp, "threads": threads, "seconds": time.perf_counter() - t}
    for key, r_ring in (("ring_0.30_as_specified", 0.3), ("ring_0.22_at_the_shock", 0.22)):
        Q = P if r_ring == 0.3 else SedovInverse(n=n, steps=steps, every=every, r_ring=r_ring)
        L, gu, gE = Q.gradient(u0, E)
        ring, target = Q.ring.numpy(), Q.target.numpy()

        def fd_of(pair):
            # (L(+h) - L(-h)) / 2h with L = sum (rho - target)^2, factored
            # cell by cell so that two nearly equal losses never cancel
            a, b = pair[0][0, ring], pair[1][0, ring]
            return float(np.sum((b - a) * (a + b - 2.0 * target))) / (2 * h)

        errs, an_all = [], []
        for (xi, xe), pair in zip(dirs, pairs):
            an = float((gu * xi * u0).sum()) + gE * float(xe * E)
            errs.append(abs(fd_of(pair) - an) / max(abs(an), 1e-300))
            an_all.append(an)
        fdE = fd_of(pairs[-1])
        rec = {"loss": L, "dL_dE": gE, "grad_u_norm": float(gu.norm()), "h": h,
               "fd_directions": len(errs), "fd_rel_err_median": float(np.median(errs)),
               "fd_rel_err_max": float(np.max(errs)),
               "fd_within_1e-6": int(np.sum(np.array(errs) <= 1e-6)),
               "E_only_fd": fdE, "E_only_fd_rel_err": abs(fdE - gE) / max(abs(gE), 1e-300)}
        if fp is not None:
            rec["branch_switches"] = {
                "directions": [dict(fp[k], fd_rel_err=errs[k]) for k in range(kfp)],
                "E_only": dict(fp[-1], fd_rel_err=rec["E_only_fd_rel_err"]),
                "directions_without_switch": int(sum(
                    1 for d in fp[:kfp] if d["steps_with_jumps"] == 0 and d["steps_with_kinks"] == 0)),
                **fp_meta}
        log(f"adjoint: {key} finite differences done")
        if args.adjoint_cpu_ref > 0:
            log(f"adjoint: {key} forward-mode CPU reference")
            import pyoracle as po
            threads = min(16, os.cpu_count() or 1)
            o = po.Oracle(Q.mesh, order=2, nthreads=threads)
            K = min(args.adjoint_cpu_ref, len(dirs))
            dc = np.zeros((K, 4, P.N))
            for k in range(K):
                xi, xe = dirs[k]
                dc[k, 1] = (Q.gamma - 1.0) * (xi * u0 + xe * Q.w).numpy()
            t = time.perf_counter()
            cT, dcT, _, _, _ = o.run_tangent(c0, dc, steps)
            resid = cT[0, ring] - target
            ref = [float(np.sum(2.0 * resid * dcT[k, 0, ring])) for k in range(K)]
            rel = [abs(an_all[k] - ref[k]) / max(abs(ref[k]), 1e-300) for k in range(K)]
            rec["cpu_reference"] = {"directions": K, "rel_err_max": float(np.max(rel)),
                                    "seconds": time.perf_counter() - t, "threads": threads,
                                    "kind": "forward-mode oracle (dual numbers)"}
        out[key] = rec
        if Q is not P:
            Q.close()
    P.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=40,
                    help="Heun steps of the CPU baseline sample, about 10 s on 16 host "
                         "threads (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cpu count)")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the per-kernel HIP-event run after the timed one")
    ap.add_argument("--profile-steps", type=int, default=50,
                    help="steps of the per-kernel HIP-event run that follows the timed run")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision run (FP32 reconstruction and HLLC)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--adjoint-n", type=int, default=256)
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--adjoint-dirs", type=int, default=32,
                    help="seeded random directions of the finite-difference check")
    ap.add_argument("--adjoint-fp-dirs", type=int, default=1, metavar="K",
                    help="directions (and E alone) whose -h/+h oracle runs are fingerprinted")
    ap.add_argument("--adjoint-cpu-ref", type=int, default=8, metavar="K",
                    help="also compare K directional derivatives with the forward-mode CPU "
                         "oracle (dual numbers, OpenMP, 8 per pass); 0 = skip")
    ap.add_argument("--no-moving", action="store_true", help="skip the moving-mesh leg")
    ap.add_argument("--moving-n", type=int, default=512)
    ap.add_argument("--moving-steps", type=int, default=200,
                    help="steps of the moving-mesh Sedov run (the full run is 200; the "
                         "plugin-level comparison runs the first 10)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ensemble-k", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (HIP runtime copy "
                         "kernel) and exit: the reference point for HBM byte counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE {world}); the multi-GPU path is a design, not a measured run.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import cases
    from rich_amd.gpu import RichGPU

    t_setup = time.perf_counter()
    mesh = cases.kh_mesh(args.nx, args.ny)
    cells = cases.kh_cells(mesh)
    N, F, G = mesh.n_cells, mesh.n_faces, mesh.n_ghosts
    nnz = int(mesh.csr_off[-1])
    g = RichGPU(mesh, order=2, cfl=0.3)
    g.set_cells(cells)
    g.init_extensives()
    ext0 = g.get_extensives()
    setup_s = time.perf_counter() - t_setup

    # CPU baseline on a bounded sample (the first steps of the same run), and
    # the GPU's state after the same steps for a full-size parity figure.
    cpu = None
    parity = None
    if args.cpu_steps > 0:
        from parity import compare
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        g.advance(args.cpu_steps, integrator=2)
        gc = g.get_cells()
        log(f"cpu baseline: {args.cpu_steps} steps on {threads} threads")
        c_ref, e_ref, dts_ref, cpu_s = cpu_baseline(mesh, cells, args.cpu_steps, threads)
        err, same = compare(gc, c_ref)
        parity = {"steps": args.cpu_steps, "max_scaled_err": err, "bit_identical": same}
        cpu = {"value": N * args.cpu_steps / cpu_s / 1e6, "unit": "Mcell-updates/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} Heun steps of the same {args.nx}x{args.ny} KH "
                         f"workload, oracle/ C restatement, OpenMP {threads} threads, "
                         f"{cpu_s:.2f} s"}
        g.set_cells(cells)
        g.init_extensives()

    log(f"fp64: {args.warmup} warm-up + {args.steps} timed steps")
    g.advance(args.warmup, integrator=2, record_dt=False)
    t0 = time.perf_counter()
    dts = g.advance(args.steps, integrator=2, record_dt=True)
    elapsed = time.perf_counter() - t0
    ext = g.get_extensives()

    value = N * args.steps / elapsed / 1e6
    c64 = g.get_cells()

    def kernel_times(ctx):
        """Per-kernel HIP-event averages over further steps (the timed run
        above records no events)."""
        if args.no_profile or args.profile_steps <= 0:
            return {}
        ctx.profile(True)
        ctx.advance(args.profile_steps, integrator=2, record_dt=False)
        out = ctx.profile_read()
        ctx.profile(False)
        return out

    prof = kernel_times(g)

    # Mixed precision: the same run (same initial state, warm-up and steps)
    # with FP32 reconstruction and HLLC and FP64 extensives; L1 error of each
    # field against the FP64 run after the same steps.
    mixed = None
    if not args.no_mixed:
        from rich_amd.gpu import RG_MIXED
        log("mixed precision run")
        gm = RichGPU(mesh, order=2, cfl=0.3, precision=RG_MIXED)
        gm.set_cells(cells)
        gm.init_extensives()
        gm.advance(args.warmup, integrator=2, record_dt=False)
        tm0 = time.perf_counter()
        gm.advance(args.steps, integrator=2, record_dt=False)
        elapsed_m = time.perf_counter() - tm0
        cm = gm.get_cells()
        prof_m = kernel_times(gm)
        vm = N * args.steps / elapsed_m / 1e6
        names = ("density", "pressure", "x_velocity", "y_velocity")
        l1 = {k: float(np.mean(np.abs(cm[j] - c64[j]))) for j, k in enumerate(names)}
        l1_rel = {k: float(np.sum(np.abs(cm[j] - c64[j])) / np.sum(np.abs(c64[j])))
                  for j, k in enumerate(names)}
        mixed = {"value": vm, "ms_per_step": 1e3 * elapsed_m / args.steps,
                 "speedup_vs_fp64": vm / value, "l1_vs_fp64": l1,
                 "l1_rel_vs_fp64": l1_rel, "steps_compared": args.warmup + args.steps}
        if prof_m:
            mixed["kernels"] = {k: {"launches": n, "avg_us": 1e3 * ms / n}
                                for k, (n, ms) in sorted(prof_m.items())}
        gm.close()
    line = {
        "metric": "Mcell-updates/s on 1 GPU (2D, 2M cells) FP64 vs mixed; %HBM roofline; adjoint/fwd",
        "value": value,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D Kelvin-Helmholtz {args.nx}x{args.ny} jittered lattice, "
                               "periodic [0,2]x[0,1], LinearGaussImproved+HLLC, Heun, static mesh",
                   "cells": N, "faces": F, "ghosts": G, "integrator": "TimeAdvance2Heun",
                   "cfl": 0.3, "gamma": 5.0 / 3.0},
    }
    b_alg = 180 * N + 48 * F
    if prof:
        # one stage = one launch each of the stage kernels (k_reduce_min:
        # the CFL minimum after the stage's update)
        stages = 2 * args.profile_steps
        stage_s = sum(prof[k][1] for k in STAGE_KERNELS if k in prof) / stages / 1e3
        traffic = [measured_traffic(k, args.nx, args.ny) for k in STAGE_KERNELS[:2]]
        achieved = b_alg / stage_s / 1e9
        line["roofline"] = {"bound": "hbm", "kernel": "stage: " + " + ".join(STAGE_KERNELS),
                            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS,
                            "traffic": sum(traffic) if None not in traffic else None,
                            "alg_bytes_per_launch": b_alg, "avg_launch_us": stage_s * 1e6,
                            "alg_bytes": "180 N + 48 F per cell-stage"}
        fused, ent, _ = g.tiles()
        if not fused:
            ent = tile_entries(mesh)
        line["kernels"] = {k: {"launches": n, "avg_us": 1e3 * ms / n,
                               "model_bytes": kernel_alg_bytes(k, N, F, G, nnz, entries=ent),
                               "model_GBs": kernel_alg_bytes(k, N, F, G, nnz, entries=ent)
                               / (ms / n / 1e3) / 1e9,
                               "traffic": measured_traffic(k, args.nx, args.ny)}
                           for k, (n, ms) in sorted(prof.items())}
        line["tile_entries"] = {"fused_stages": fused, "entries": ent, "per_face": ent / F,
                                "cells_per_tile": TILE}
    stage_gbs = b_alg * 2 * args.steps / elapsed / 1e9
    line["stage_roofline"] = {"bytes_per_stage": b_alg, "achieved": stage_gbs,
                              "frac": stage_gbs / HBM_PEAK_GBS, "unit": "GB/s"}
    if mixed:
        line["mixed"] = mixed
    if not args.no_moving:
        line["moving_mesh"] = moving_leg(args)
    if not args.no_ensemble:
        line["ensemble"] = ensemble_leg(args)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_leg(args)
    if cpu:
        line["cpu_baseline"] = cpu
        line["speedup_vs_cpu"] = value / cpu["value"]
    if parity:
        line["parity_vs_oracle"] = parity
    line["conservation"] = {
        "mass": float(ext[0].sum() / ext0[0].sum() - 1.0),
        "energy": float(ext[1].sum() / ext0[1].sum() - 1.0),
        "momentum_x": float(ext[2].sum() - ext0[2].sum()),
        "momentum_y": float(ext[3].sum() - ext0[3].sum()),
    }
    line["sim_time"] = float(np.sum(dts))
    line["setup_s"] = setup_s
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
