"""Benchmark: lattice site-updates/s of the 2D cell-sorting sweep on one GPU.

Workload: 1024^2 lattice, 63x63 = 3969 cells of two types seeded as 10x10
squares tiling the centre; area target 100 (lambda 25), perimeter target 40
(lambda 2), J_mA = J_mB = 16, J_AA = 2, J_BB = 14, J_AB = 11, T = 10.
A step is one ``run()`` of ``--mcs-per-step`` Monte Carlo steps (MCS); one
MCS is L^2 copy attempts (site-updates).  The lattice and cell tables are
resident in HBM before timing starts.

Prints one JSON line with the throughput, the HBM roofline of the sweep
kernel (8 algorithmic bytes per site-update: one int32 cell-id read and one
write) from HIP events on the engine's stream, and the CPU oracle (same
sweep, same schedule, OpenMP over the host cores) timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-cpm_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mcs-per-step", type=int, default=200)
    ap.add_argument("--workload", default="sorting2d",
                    choices=["sorting2d", "plain3d", "plain2d", "batched"])
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--checkerboard", type=int, default=1, help="positions_per_checkerboard c")
    ap.add_argument("--updates", type=int, default=1, help="updates_per_checkerboard_switch u")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (hipMemcpy), the known-size "
                         "reference for rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def cpu_baseline(wl, c, u, seconds):
    """Time the CPU oracle (OpenMP over all host cores) on a bounded sample of
    the same workload, from the same initial lattice."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O
    p = wl.params()
    prm = O.Params.make(p["temperature"], J=p["J"], S=p["S"], lam_area=p["lam_area"],
                        target_area=p["target_area"], lam_perim=p["lam_perim"],
                        target_perim=p["target_perim"])
    cores = len(os.sched_getaffinity(0))
    lat = wl.lattice.copy()
    tab = O.tally(lat, wl.d, wl.ncells)
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 0, 2, cores)  # calibration
    per_mcs = max((time.perf_counter() - t) / 2, 1e-6)
    n = max(2, int(seconds / per_mcs))
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 2, n, cores)
    dt = time.perf_counter() - t
    sites = wl.L ** wl.d
    return {"value": n * sites / dt, "unit": "site-updates/s", "cores": cores, "kind": "port",
            "sample": f"{n} MCS of the same {wl.L}^{wl.d} workload (MCS 2..{n + 1}) from its "
                      f"initial lattice, oracle/libcpm_oracle.so with OpenMP, {dt:.1f} s",
            "mcs_per_s": n / dt}


def bench_batched(a, log):
    """Batched replicas: 64 sorting lattices of mixed sizes and parameters in
    one batch (one launch per colour phase for all of them) vs the same
    replicas run one after another; both must end bit-identical."""
    import hashlib

    import numpy as np

    import gpucpm
    from gpucpm import workloads

    wls = workloads.batched_sorting(nrep=a.replicas, seed=4)
    c, u = a.checkerboard, a.updates
    mcs = a.steps * a.mcs_per_step
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=2 * c, positions_per_checkerboard=c,
              updates_per_checkerboard_switch=u, updates_per_barrier=1)
    sites = sum(w.L ** 2 for w in wls)

    def fresh():
        sims = []
        for i, wl in enumerate(wls):
            s = wl.make_cpm()
            s.push_to_gpu()
            s.set_seed(9000 + i)
            sims.append(s)
        return sims

    def digest(sims):
        h = hashlib.sha256()
        for s in sims:
            s.pull_from_gpu()
            h.update(s.get_state().tobytes())
        return h.hexdigest()

    log(f"batched: {len(wls)} replicas, {sites / 1e6:.1f} M sites, {mcs} MCS; batch")
    sims = fresh()
    batch = gpucpm.Batch(sims)
    batch.run(**kw, iterations=a.mcs_per_step, inner_iterations=a.mcs_per_step)  # warmup
    batch.synchronize()
    sims = fresh()
    batch = gpucpm.Batch(sims)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        batch.run(**kw, iterations=a.mcs_per_step, inner_iterations=a.mcs_per_step)
    batch.synchronize()
    t_batch = time.perf_counter() - t0
    d_batch = digest(sims)
    log(f"batch {t_batch:.3f} s; sequential")
    sims = fresh()
    t0 = time.perf_counter()
    for s in sims:
        for _ in range(a.steps):
            s.run(**kw, iterations=a.mcs_per_step, inner_iterations=a.mcs_per_step)
        s.synchronize()
    t_seq = time.perf_counter() - t0
    d_seq = digest(sims)
    log(f"sequential {t_seq:.3f} s")
    out = {
        "metric": "lattice site-updates/sec (batched replicas)",
        "value": mcs * sites / t_batch,
        "unit": "site-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": 1,
        "ms_per_step": t_batch / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{len(wls)} sorting replicas, L in {{128,256,512,1024}}, "
                               "J_AB in [6,14], T in [5,20], L^2/250 cells",
                   "replicas": len(wls), "sites": sites, "mcs": mcs,
                   "positions_per_checkerboard": c, "updates_per_checkerboard_switch": u},
        "sequential_value": mcs * sites / t_seq,
        "speedup_vs_sequential": t_seq / t_batch,
        "bit_identical_to_sequential": d_batch == d_seq,
    }
    print(json.dumps(out), flush=True)


def copy_calibration(mib):
    """Three device-to-device copies of a known size through the HIP runtime
    (its copy kernel reads and writes exactly mib MiB per copy)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(replicas shard across GPUs without exchange; see the sharding design).",
              file=sys.stderr)
        sys.exit(2)
    log = lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True)
    if a.workload == "batched":
        return bench_batched(a, log)
    import gpucpm
    from gpucpm import workloads

    wl = {"sorting2d": workloads.sorting2d, "plain3d": workloads.plain3d,
          "plain2d": workloads.plain2d}[a.workload]()
    c, u = a.checkerboard, a.updates
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(20240601)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=2 * c, positions_per_checkerboard=c,
              updates_per_checkerboard_switch=u, updates_per_barrier=1,
              iterations=a.mcs_per_step, inner_iterations=a.mcs_per_step)
    log(f"{a.workload} L={wl.L} d={wl.d} cells={wl.ncells} c={c} u={u} "
        f"mode={os.environ.get('GPUCPM_MODE', 'default')}: warmup {a.warmup} x {a.mcs_per_step} MCS")
    for _ in range(a.warmup):
        sim.run(**kw)
    sim.synchronize()
    log(f"timed: {a.steps} x {a.mcs_per_step} MCS")

    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.run(**kw)
    sim.synchronize()
    wall = time.perf_counter() - t0

    log(f"timed region {wall:.3f} s; kernel timing pass")
    # Kernel durations of the same number of steps, bracketed per launch.
    sim.timing_begin()
    for _ in range(a.steps):
        sim.run(**kw)
    kernel_ms, launches, updates = sim.timing_end()
    launch = sim.last_launch()

    sites = wl.L ** wl.d
    mcs = a.steps * a.mcs_per_step
    value = mcs * sites / wall
    avg_launch_s = kernel_ms / 1e3 / launches
    bytes_per_launch = updates / launches * BYTES_PER_UPDATE
    achieved = bytes_per_launch / avg_launch_s / 1e9
    workload_desc = {
        "sorting2d": "2D cell sorting 1024^2, 3969 cells (10x10 squares), two types, 10000-MCS config",
        "plain3d": "3D CPM 256^3, 20000 cubes 5^3, 26-neighbourhood",
        "plain2d": "plain 2D CPM 256^2, 100 cells",
    }[a.workload]
    out = {
        "metric": "lattice site-updates/sec",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": workload_desc, "L": wl.L, "d": wl.d, "cells": wl.ncells,
                   "mcs_per_step": a.mcs_per_step, "positions_per_checkerboard": c,
                   "updates_per_checkerboard_switch": u, "seed": 20240601},
        "mcs_per_s": mcs / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": launch["kernel"], "avg_launch_us": avg_launch_s * 1e6,
                     "bytes_per_launch": bytes_per_launch, "launches": launches,
                     "grid_blocks": launch["blocks"], "lds_tile": launch["tile"]},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        log(f"CPU oracle sample (~{a.cpu_seconds:.0f} s)")
        out["cpu_baseline"] = cpu_baseline(wl, c, u, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
