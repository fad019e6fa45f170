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
import re
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
                    choices=["sorting2d", "plain3d", "plain2d", "batched", "gradient", "interactive"])
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--checkerboard", type=int, default=1, help="positions_per_checkerboard c")
    ap.add_argument("--updates", type=int, default=1, help="updates_per_checkerboard_switch u")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="let the engine time its kernel forms during warmup and keep the fastest")
    ap.add_argument("--no-grad", action="store_true", help="skip the dH/dtheta check")
    ap.add_argument("--pmc-summary", nargs=6, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR",
                                                       "COPY_WRITE_DIR", "KERNEL", "OUT_JSON"),
                    help="summarize rocprofv3 FETCH_SIZE / WRITE_SIZE runs of KERNEL, scaled by the "
                         "same counters on a --copy-calibration run, into OUT_JSON")
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
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 0, 1, cores)  # thread pool start-up
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 1, 2, cores)  # calibration
    per_mcs = max((time.perf_counter() - t) / 2, 1e-6)
    n = max(2, int(seconds / per_mcs))
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 3, n, cores)
    dt = time.perf_counter() - t
    sites = wl.L ** wl.d
    # The same sweep on one core, continuing from where the sample ended.
    m = n + 3
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, m, 1, 1)
    per_mcs1 = max(time.perf_counter() - t, 1e-6)
    n1 = max(1, int(seconds / 3 / per_mcs1))
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, m + 1, n1, 1)
    dt1 = time.perf_counter() - t
    return {"value": n * sites / dt, "unit": "site-updates/s", "cores": cores, "kind": "port",
            "sample": f"{n} MCS of the same {wl.L}^{wl.d} workload (MCS 3..{n + 2}) from its "
                      f"initial lattice, oracle/libcpm_oracle.so with OpenMP, {dt:.1f} s",
            "mcs_per_s": n / dt,
            "single_core": {"value": n1 * sites / dt1, "mcs_per_s": n1 / dt1,
                            "sample": f"{n1} MCS (MCS {m + 1}..{m + n1}), 1 thread, {dt1:.1f} s"}}


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
    # shared = 1: the reference's shared-memory schedule (K = 2 colours)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=2 * c, positions_per_checkerboard=c,
              updates_per_checkerboard_switch=u, updates_per_barrier=1, shared=1)
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
    batch.set_autotune(a.autotune)  # (its choices carry over to the timed batch)
    batch.run(**kw, iterations=a.mcs_per_step, inner_iterations=a.mcs_per_step)  # warmup
    batch.synchronize()
    sims = fresh()
    batch = gpucpm.Batch(sims)
    batch.set_autotune(a.autotune)
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
                   "positions_per_checkerboard": c, "updates_per_checkerboard_switch": u,
                   "autotune": bool(a.autotune)},
        "sequential_value": mcs * sites / t_seq,
        "speedup_vs_sequential": t_seq / t_batch,
        "bit_identical_to_sequential": d_batch == d_seq,
    }
    print(json.dumps(out), flush=True)


def _oracle_energy_fn(lats, wls, layout):
    """FP64 H(theta) averaged over lattices, from the oracle's own reduction."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O
    stats = []
    for lat, wl in zip(lats, wls):
        st = O.energy_stats(lat, 2, wl.ncells, O.Params.make(1.0))  # targets 0: raw moments
        spin, typ = lat & 0xFFFFFF, lat >> 24
        ncell = [len(np.unique(spin[(typ == t) & (spin > 0)])) for t in range(layout.nt)]
        stats.append((st, ncell))

    def H(th):
        nt, nj = layout.nt, layout.nj
        J = np.zeros((20, 20))
        for k, (a, b) in enumerate(layout.pairs):
            J[a, b] = th[k]
        tot = 0.0
        for st, ncell in stats:
            tot += float(np.sum(J * st["pairs"]))
            for t in range(nt):
                la, a0 = th[nj + t], th[nj + nt + t]
                lp, p0 = th[nj + 2 * nt + t], th[nj + 3 * nt + t]
                tot += la * (st["sa2"][t] - 2 * a0 * st["sa1"][t] + ncell[t] * a0 * a0)
                tot += lp * (st["sp2"][t] - 2 * p0 * st["sp1"][t] + ncell[t] * p0 * p0)
        return tot / len(stats)
    return H


def gradient_check(nrep, L, mcs, log, J_AB=11, seed0=0, contrast=False):
    """Replicas of the sorting lattice (C2 parameters) run forward mcs MCS as
    one batch; then the backward statistics kernel gives dH/dtheta (FP32)
    for J, lambda_A, A0, lambda_P, P0, checked against central differences
    (step 1e-3) of the oracle's FP64 Hamiltonian on the same lattices."""
    import numpy as np
    import torch

    import gpucpm
    from gpucpm import workloads
    from gpucpm.autograd import ThetaLayout, contrastive_gradient, energy

    grid = (L - 112) // 10
    layout = ThetaLayout(3)

    def make(J, seed):
        wls = [workloads.sorting2d(L=L, grid=grid, edge=10, seed=seed + i, J_AB=J) for i in range(nrep)]
        sims = []
        for i, wl in enumerate(wls):
            s = wl.make_cpm()
            s.push_to_gpu()
            s.set_seed(7000 + seed + i)
            sims.append(s)
        return wls, sims

    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4, positions_per_thread=2,
              positions_per_checkerboard=1, updates_per_checkerboard_switch=1, updates_per_barrier=1,
              iterations=mcs, inner_iterations=mcs, shared=1)
    wls, sims = make(J_AB, seed0)
    log(f"gradient: {nrep} x {L}^2 forward {mcs} MCS")
    t0 = time.perf_counter()
    gpucpm.Batch(sims).run(**kw)
    for s in sims:
        s.synchronize()
    t_fwd = time.perf_counter() - t0
    p = wls[0].params()
    theta = layout.pack(p["J"], p["lam_area"], p["target_area"], p["lam_perim"], p["target_perim"],
                        dtype=torch.float32).requires_grad_(True)
    t0 = time.perf_counter()
    H = energy(theta, sims, layout)
    H.backward()
    t_bwd = time.perf_counter() - t0
    g = theta.grad.double().numpy()
    lats = []
    for s in sims:
        s.pull_from_gpu()
        lats.append(s.get_state().view(np.uint32).copy())
    H64 = _oracle_energy_fn(lats, wls, layout)
    th = theta.detach().double().numpy()
    h = 1e-3
    fd = np.array([(H64(th + h * e) - H64(th - h * e)) / (2 * h) for e in np.eye(th.size)])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1.0)
    out = {"gradient_rel_error": float(rel.max()),
           "energy_rel_error": abs(float(H.detach()) - H64(th)) / abs(H64(th)),
           "replicas": nrep, "L": L, "mcs": mcs, "forward_s": t_fwd,
           "backward_s": t_bwd, "backward_sites_per_s": nrep * L * L / t_bwd}
    if contrast:
        log(f"contrastive: data sampled at J_AB = 8")
        _, data = make(8, seed0 + 1000)
        gpucpm.Batch(data).run(**kw)
        for s in data:
            s.synchronize()
        gc = contrastive_gradient(theta, data, sims, layout).numpy()
        k = layout.pairs.index((1, 2))
        out["contrastive_grad_J_AB"] = float(gc[k])
        out["contrastive_step_moves_J_AB_toward_8"] = bool(gc[k] > 0)  # theta -= lr * grad lowers J_AB
    return out


def bench_interactive(a, log):
    """The notebook pattern (examples/gpucpm_sorting_example.ipynb: one
    run(iterations=ticks, inner_iterations=1) per frame, then a readback):
    frames per second on the 256^2 sorting lattice with one MCS and one
    pull_from_gpu per frame, the default shared=0 roles of the notebook
    (positions_per_thread 4, c 2)."""
    import gpucpm
    from gpucpm import workloads
    wl = workloads.sorting2d(L=256, grid=20, edge=10, seed=2)
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(20240601)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=4, positions_per_checkerboard=2,
              updates_per_checkerboard_switch=1, updates_per_barrier=1,
              iterations=1, inner_iterations=1, shared=0, partial_dispatch=1)
    frames = a.steps * a.mcs_per_step
    for _ in range(a.warmup * 10):
        sim.run(**kw)
        sim.pull_from_gpu()
    log(f"interactive: {frames} frames of 1 MCS + readback at 256^2")
    t0 = time.perf_counter()
    for _ in range(frames):
        sim.run(**kw)
        sim.pull_from_gpu()
    wall = time.perf_counter() - t0
    sim.timing_begin()
    for _ in range(a.steps):
        sim.run(**kw)
    kernel_ms, launches, updates = sim.timing_end()
    launch = sim.last_launch()
    return {
        "metric": "frames/s (1 MCS + lattice readback per frame)",
        "value": frames / wall,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "interactive notebook pattern: 2D sorting 256^2, 400 cells, "
                               "run(iterations=1) + pull per frame, shared=0, ppt 4, c 2",
                   "frames": frames},
        "ms_per_frame": wall / frames * 1e3,
        "kernel_ms_per_mcs": kernel_ms / a.steps,
        "launches_per_mcs": launches / a.steps,
        "kernel": launch["kernel"],
        "roofline": None,
        "cpu_baseline": None,
    }


def bench_gradient(a, log):
    out = gradient_check(32, 512, a.steps * a.mcs_per_step, log, contrast=True)
    line = {
        "metric": "gradient rel. error (dH/dtheta FP32 vs FP64 central differences)",
        "value": out["gradient_rel_error"], "unit": "relative", "n_gpus": 1,
        "steps": a.steps, "warmup": 0, "ms_per_step": out["forward_s"] / a.steps * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32 sweep, int64 statistics, fp32 gradient", "data": "synthetic",
        "config": {"workload": "32 replicas of 512^2 two-type sorting (C2 parameters), forward then "
                               "dH/dtheta for J (3x3), lambda_A, A0, lambda_P, P0; FD step 1e-3; "
                               "contrastive step toward J_AB = 8"},
        "roofline": {"bound": "hbm", "kernel": "k_stats_pairs + k_stats_cells",
                     "achieved": out["backward_sites_per_s"] * 4 / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": out["backward_sites_per_s"] * 4 / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "note": "backward time includes the host side: one wait for all replicas' statistics and the per-replica energy and gradient from them"},
        "cpu_baseline": None,
    }
    line.update({k: v for k, v in out.items() if k != "gradient_rel_error"})
    print(json.dumps(line), flush=True)


def kernel_token(name):
    """File-name form of a kernel name: 'k_sweep_phase_compact<2, true, false>'
    -> 'k_sweep_phase_compact_2_true_false'."""
    return re.sub(r"[^A-Za-z0-9]+", "_", name).strip("_")


def _pmc_mean(d, kernel):
    import csv
    import glob
    vals = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if kernel in r["Kernel_Name"]:
                vals.append(float(r["Counter_Value"]))
    if not vals:
        raise SystemExit(f"no {kernel} dispatches in {d}")
    return sum(vals) / len(vals), len(vals)


def pmc_summary(fetch_dir, write_dir, cfetch_dir, cwrite_dir, kernel, out):
    """HBM bytes per dispatch of KERNEL from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch), each scaled so that the runtime's copy kernel of a known size
    reads and writes exactly that size."""
    f, nf = _pmc_mean(fetch_dir, kernel)
    w, nw = _pmc_mean(write_dir, kernel)
    cf, _ = _pmc_mean(cfetch_dir, "copyBuffer")
    cw, _ = _pmc_mean(cwrite_dir, "copyBuffer")
    true_kib = 256 * 1024  # --copy-calibration 256
    kf, kw = true_kib / cf, true_kib / cw
    res = {"kernel": kernel, "fetch_kib": f, "write_kib": w, "dispatches": [nf, nw],
           "copy_fetch_kib": cf, "copy_write_kib": cw, "fetch_scale": kf, "write_scale": kw,
           "bytes_per_dispatch": (f * kf + w * kw) * 1024}
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res))


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
    if a.pmc_summary:
        return pmc_summary(*a.pmc_summary)
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(replicas shard across GPUs without exchange; see the sharding design).",
              file=sys.stderr)
        sys.exit(2)
    log = lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True)
    if a.workload == "batched":
        return bench_batched(a, log)
    if a.workload == "interactive":
        out = bench_interactive(a, log)
        print(json.dumps(out), flush=True)
        return 0
    if a.workload == "gradient":
        return bench_gradient(a, log)
    import gpucpm
    from gpucpm import workloads

    wl = {"sorting2d": workloads.sorting2d, "plain3d": workloads.plain3d,
          "plain2d": workloads.plain2d}[a.workload]()
    c, u = a.checkerboard, a.updates
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(20240601)
    sim.set_autotune(a.autotune)
    # shared = 1: the reference's shared-memory schedule (K = 2 colours)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=2 * c, positions_per_checkerboard=c,
              updates_per_checkerboard_switch=u, updates_per_barrier=1,
              iterations=a.mcs_per_step, inner_iterations=a.mcs_per_step, shared=1)
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
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{kernel_token(launch['kernel'])}_{a.workload}.json")
    if os.path.exists(tfile) and c == 1 and u == 1:
        traffic = json.load(open(tfile))["bytes_per_dispatch"]
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
                   "updates_per_checkerboard_switch": u, "seed": 20240601,
                   "autotune": bool(a.autotune)},
        "mcs_per_s": mcs / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": launch["kernel"], "avg_launch_us": avg_launch_s * 1e6,
                     "bytes_per_launch": bytes_per_launch, "launches": launches,
                     "grid_blocks": launch["blocks"], "lds_tile": launch["tile"]},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        log(f"CPU oracle sample (~{a.cpu_seconds:.0f} s)")
        out["cpu_baseline"] = cpu_baseline(wl, c, u, a.cpu_seconds)
    if a.workload == "sorting2d" and not a.no_grad:
        g = gradient_check(4, 512, 200, log)
        out["gradient_rel_error"] = g["gradient_rel_error"]
        out["gradient_check"] = {k: g[k] for k in ("replicas", "L", "mcs", "energy_rel_error")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
