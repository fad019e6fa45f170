"""Benchmark: lattice site-updates/s of the 2D cell-sorting sweep on one GPU.

Workload (the headline record): 1024^2 lattice, 63x63 = 3969 cells of two
types seeded as 10x10 squares tiling the centre; area target 100 (lambda
25), perimeter target 40 (lambda 2), J_mA = J_mB = 16, J_AA = 2, J_BB = 14,
J_AB = 11, T = 10.  A step is one ``run()`` of ``--mcs-per-step`` Monte
Carlo steps (MCS); one MCS is L^2 copy attempts (site-updates).  The
lattice and cell tables are resident in HBM before timing starts.

Prints one JSON line with the throughput, the HBM roofline of the sweep
kernel (8 algorithmic bytes per site-update: one int32 cell-id read and one
write) from HIP events on the engine's stream, and the CPU oracle (same
sweep, same schedule, OpenMP over the host cores) timed on a bounded
sample.  With the default workload the line also carries ``configs``, one
record per other configuration, each with its own roofline and baseline:
3D 256^3 (20,000 cells), 64 batched replicas over 2,000 MCS (against the same
replicas run one after another and run concurrently on their own streams),
and the dH/dtheta backward of 32 x 512^2 replicas after 2,000 MCS (with the
gradient error against FP64 central differences and a contrastive fit).
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
    ap.add_argument("--no-grad", action="store_true", help="skip the contrastive fit in the gradient record")
    ap.add_argument("--launch", type=int, default=-1,
                    help="launch form (-1 auto, 0 per phase, 3 one workgroup, 4 resident 2D tiles)")
    ap.add_argument("--batch-only", action="store_true",
                    help="batched: only the batch runs (no sequential, concurrent, graph or form records), "
                         "for profiling the batched kernel on the C4 shape alone")
    ap.add_argument("--no-configs", action="store_true",
                    help="sorting2d: only the headline record (no C3 / C4 / C5 sub-records)")
    ap.add_argument("--pmc-summary", nargs=6, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR",
                                                       "COPY_WRITE_DIR", "KERNEL", "OUT_JSON"),
                    help="summarize rocprofv3 FETCH_SIZE / WRITE_SIZE runs of KERNEL, scaled by the "
                         "same counters on a --copy-calibration run, into OUT_JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (hipMemcpy), the known-size "
                         "reference for rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def cpu_baseline(wl, c, u, seconds, shared=True, ppt=None):
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
    sk = dict(shared=shared, ppt=ppt)
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 0, 1, cores, **sk)  # thread pool start-up
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 1, 2, cores, **sk)  # calibration
    per_mcs = max((time.perf_counter() - t) / 2, 1e-6)
    n = max(2, int(seconds / per_mcs))
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, 3, n, cores, **sk)
    dt = time.perf_counter() - t
    sites = wl.L ** wl.d
    # The same sweep on one core, continuing from where the sample ended.
    m = n + 3
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, m, 1, 1, **sk)
    per_mcs1 = max(time.perf_counter() - t, 1e-6)
    n1 = max(1, int(seconds / 3 / per_mcs1))
    t = time.perf_counter()
    O.sweep(lat, wl.d, tab, prm, c, u, 1, m + 1, n1, 1, **sk)
    dt1 = time.perf_counter() - t
    return {"value": n * sites / dt, "unit": "site-updates/s", "cores": cores, "kind": "port",
            "sample": f"{n} MCS of the same {wl.L}^{wl.d} workload (MCS 3..{n + 2}) from its "
                      f"initial lattice, oracle/libcpm_oracle.so with OpenMP, {dt:.1f} s",
            "mcs_per_s": n / dt,
            "single_core": {"value": n1 * sites / dt1, "mcs_per_s": n1 / dt1,
                            "sample": f"{n1} MCS (MCS {m + 1}..{m + n1}), 1 thread, {dt1:.1f} s"}}


def storage_of(kernel):
    """Site storage of the timed kernel: the sweep runs on a lossless 16-bit
    re-encoding of the int32 lattice when every site fits."""
    if "unsigned short" in kernel:
        return "uint16 sites (cell | type << b, b = 16 - type bits): lossless re-encoding of the int32 lattice"
    return "uint32 sites (cell | type << 24)"


def measure_single(wl, name, c, u, steps, warmup, mcs_per_step, cpu_seconds, no_cpu, autotune, log,
                   launch=-1, shared=1, ppt=None):
    """Whole-run throughput of one engine (inputs resident before timing),
    the sweep kernel's roofline from per-launch HIP events, and the CPU
    oracle on a bounded sample of the same workload."""
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(20240601)
    sim.set_autotune(autotune)
    sim.set_launch(launch)
    # shared = 1: the reference's shared-memory schedule (K = 2 colours);
    # shared = 0: K = ppt / c colours (the reference's default)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=ppt or 2 * c, positions_per_checkerboard=c,
              updates_per_checkerboard_switch=u, updates_per_barrier=1,
              iterations=mcs_per_step, inner_iterations=mcs_per_step, shared=shared)
    log(f"{name} L={wl.L} d={wl.d} cells={wl.ncells} c={c} u={u}: warmup {warmup} x {mcs_per_step} MCS")
    for _ in range(warmup):
        sim.run(**kw)
    sim.synchronize()
    log(f"{name} timed: {steps} x {mcs_per_step} MCS")
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.run(**kw)
    sim.synchronize()
    wall = time.perf_counter() - t0
    log(f"{name} timed region {wall:.3f} s; kernel timing pass")
    # Kernel durations of the same number of steps, bracketed per launch.
    sim.timing_begin()
    for _ in range(steps):
        sim.run(**kw)
    kernel_ms, launches, updates = sim.timing_end()
    launch = sim.last_launch()
    sites = wl.L ** wl.d
    mcs = steps * mcs_per_step
    avg_launch_s = kernel_ms / 1e3 / launches
    bytes_per_launch = updates / launches * BYTES_PER_UPDATE
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = _profiled_traffic(launch["kernel"], name) if c == 1 and u == 1 and shared else None
    out = {
        "value": mcs * sites / wall,
        "unit": "site-updates/s",
        "ms_per_step": wall / steps * 1e3,
        "mcs_per_s": mcs / wall,
        "storage": storage_of(launch["kernel"]),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": launch["kernel"], "avg_launch_us": avg_launch_s * 1e6,
                     "bytes_per_launch": bytes_per_launch, "launches": launches,
                     "grid_blocks": launch["blocks"], "lds_tile": launch["tile"]},
        "cpu_baseline": None,
    }
    if not no_cpu:
        log(f"{name} CPU oracle sample (~{cpu_seconds:.0f} s)")
        out["cpu_baseline"] = cpu_baseline(wl, c, u, cpu_seconds, shared=bool(shared), ppt=ppt)
    return out


def measure_batched(nrep, mcs, log, autotune=False, batch_only=False):
    """Batched replicas (64 sorting lattices of mixed sizes and parameters)
    in one batch, one launch per colour phase for all of them, against the
    same replicas run one after another (each run waited for) and run
    concurrently on their own engines' streams (one wait at the end); all
    three must end bit-identical."""
    import hashlib

    import gpucpm
    from gpucpm import workloads

    wls = workloads.batched_sorting(nrep=nrep, seed=4)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=2, positions_per_checkerboard=1,
              updates_per_checkerboard_switch=1, updates_per_barrier=1, shared=1)
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

    log(f"batched: {len(wls)} replicas, {sites / 1e6:.1f} M sites, {mcs} MCS; warmup")
    batch = gpucpm.Batch(fresh())
    batch.set_autotune(autotune)  # (its choices carry over to the timed batch)
    batch.run(**kw, iterations=100, inner_iterations=100)
    batch.synchronize()
    sims = fresh()
    batch = gpucpm.Batch(sims)
    batch.set_autotune(autotune)
    t0 = time.perf_counter()
    batch.run(**kw, iterations=mcs, inner_iterations=mcs)
    batch.synchronize()
    t_batch = time.perf_counter() - t0
    d_batch = digest(sims)
    log(f"batch {t_batch:.3f} s; kernel timing pass")
    tsims = fresh()
    tb = gpucpm.Batch(tsims)
    tb.timing_begin()
    tb.run(**kw, iterations=200, inner_iterations=200)
    kernel_ms, launches, updates = tb.timing_end()
    kernel = tb.last_kernel
    if batch_only:
        avg_launch_s = kernel_ms / 1e3 / launches
        bytes_per_launch = updates / launches * BYTES_PER_UPDATE
        achieved = bytes_per_launch / avg_launch_s / 1e9
        return {"workload": f"{len(wls)} sorting replicas (batch runs only)",
                "value": mcs * sites / t_batch, "unit": "site-updates/s", "ms_total": t_batch * 1e3,
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel,
                             "avg_launch_us": avg_launch_s * 1e6, "bytes_per_launch": bytes_per_launch,
                             "launches": launches}}
    log("graph replay")
    gsims = fresh()
    gb = gpucpm.Batch(gsims)
    gb.set_form(2)
    t0 = time.perf_counter()
    gb.run(**kw, iterations=mcs, inner_iterations=mcs)  # (the capture is inside the timed run)
    gb.synchronize()
    t_graph = time.perf_counter() - t0
    d_graph = digest(gsims)
    graph_info = gb.graph_info()
    log("sequential runs")
    sims = fresh()
    t0 = time.perf_counter()
    for s in sims:
        s.run(**kw, iterations=mcs, inner_iterations=mcs)
        s.synchronize()
    t_seq = time.perf_counter() - t0
    d_seq = digest(sims)
    log("concurrent runs on the engines' streams")
    sims = fresh()
    t0 = time.perf_counter()
    for s in sims:
        s.run(**kw, iterations=mcs, inner_iterations=mcs)
    for s in sims:
        s.synchronize()
    t_conc = time.perf_counter() - t0
    d_conc = digest(sims)
    log("batch forms on a launch-bound batch")
    forms_small = _batch_forms_small(kw)
    avg_launch_s = kernel_ms / 1e3 / launches
    bytes_per_launch = updates / launches * BYTES_PER_UPDATE
    achieved = bytes_per_launch / avg_launch_s / 1e9
    return {
        "workload": f"{len(wls)} sorting replicas, L in {{128,256,512,1024}}, J_AB in [6,14], "
                    "T in [5,20], L^2/250 cells, 2000-MCS config",
        "value": mcs * sites / t_batch, "unit": "site-updates/s", "replicas": len(wls),
        "sites": sites, "mcs": mcs, "ms_total": t_batch * 1e3,
        "storage": storage_of(kernel),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _profiled_traffic(kernel, "batched"),
                     "kernel": kernel, "avg_launch_us": avg_launch_s * 1e6, "bytes_per_launch": bytes_per_launch,
                     "launches": launches},
        "sequential": {"value": mcs * sites / t_seq, "ms_total": t_seq * 1e3,
                       "note": "each replica's run() waited for before the next"},
        "concurrent_streams": {"value": mcs * sites / t_conc, "ms_total": t_conc * 1e3,
                               "note": "every replica's run() queued on its own stream, one wait"},
        "graph_replay": {"value": mcs * sites / t_graph, "ms_total": t_graph * 1e3,
                         "form_used": graph_info[0], "captures": graph_info[1], "replays": graph_info[2],
                         "note": f"batch form 2: the per-phase launches of {mcs // max(graph_info[2], 1)} MCS "
                                 f"captured once (inside the timed run) and replayed {graph_info[2]} times"},
        "forms_small_batch": forms_small,
        "speedup_vs_sequential": t_seq / t_batch,
        "speedup_vs_concurrent_streams": t_conc / t_batch,
        "bit_identical": d_batch == d_seq == d_conc == d_graph,
        "cpu_baseline": None,
    }


def _batch_forms_small(kw, nrep=16, L=64, mcs=1280):
    """MCS/s of a launch-bound batch (16 sorting replicas of 64^2) in each
    batch form: per-phase launches, one persistent launch per run, hip graph
    replay (captured in an untimed warm-up run of the same shape)."""
    import gpucpm
    from gpucpm import workloads
    wls = workloads.batched_sorting(nrep=nrep, seed=4, sizes=(L,))
    out = {"workload": f"{nrep} sorting replicas of {L}^2, {mcs} MCS per run"}
    for name, form in (("per_phase", 0), ("persistent", 1), ("graph_replay", 2)):
        sims = []
        for i, wl in enumerate(wls):
            s = wl.make_cpm()
            s.push_to_gpu()
            s.set_seed(9 + i)
            sims.append(s)
        b = gpucpm.Batch(sims)
        b.set_form(form)
        b.run(**kw, iterations=64, inner_iterations=64)
        b.synchronize()
        t0 = time.perf_counter()
        b.run(**kw, iterations=mcs, inner_iterations=mcs)
        b.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"mcs_per_s": mcs / dt, "us_per_mcs": dt / mcs * 1e6}
    return out


def _oracle_energy_fn(lats, wls, layout):
    """FP64 H(theta) averaged over lattices, from the oracle's own reduction;
    also the seconds the oracle took to reduce the lattices."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O
    stats = []
    t0 = time.perf_counter()
    for lat, wl in zip(lats, wls):
        st = O.energy_stats(lat, 2, wl.ncells, O.Params.make(1.0))  # targets 0: raw moments
        stats.append(st)
    t_oracle = time.perf_counter() - t0
    for st, lat in zip(stats, lats):
        spin, typ = lat & 0xFFFFFF, lat >> 24
        st["ncell"] = [len(np.unique(spin[(typ == t) & (spin > 0)])) for t in range(layout.nt)]

    def H(th):
        nt, nj = layout.nt, layout.nj
        J = np.zeros((20, 20))
        for k, (a, b) in enumerate(layout.pairs):
            J[a, b] = th[k]
        tot = 0.0
        for st in stats:
            tot += float(np.sum(J * st["pairs"]))
            for t in range(nt):
                la, a0 = th[nj + t], th[nj + nt + t]
                lp, p0 = th[nj + 2 * nt + t], th[nj + 3 * nt + t]
                n = st["ncell"][t]
                tot += la * (st["sa2"][t] - 2 * a0 * st["sa1"][t] + n * a0 * a0)
                tot += lp * (st["sp2"][t] - 2 * p0 * st["sp1"][t] + n * p0 * p0)
        return tot / len(stats)
    return H, t_oracle


def measure_gradient(nrep, L, mcs, log, fit=True):
    """Replicas of the sorting lattice (C2 parameters) run forward mcs MCS as
    one batch; the batched statistics kernel then gives dH/dtheta (FP32) for
    J, lambda_A, A0, lambda_P, P0, checked against central differences (step
    1e-3) of the oracle's FP64 Hamiltonian on the same lattices.  The fit:
    J_12 from 11 toward data sampled at J_12 = 8 (gpucpm.fit), on the
    reference sorting example's Hamiltonian (with the C2 parameters the
    interfaces do not move and J_12 leaves no trace in the samples)."""
    import numpy as np
    import torch

    import gpucpm
    from gpucpm import workloads
    from gpucpm.autograd import ThetaLayout, energy

    grid = (L - 112) // 10
    layout = ThetaLayout(3)
    wls = [workloads.sorting2d(L=L, grid=grid, edge=10, seed=i) for i in range(nrep)]
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(7000 + i)
        sims.append(s)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4, positions_per_thread=2,
              positions_per_checkerboard=1, updates_per_checkerboard_switch=1, updates_per_barrier=1,
              shared=1)
    log(f"gradient: {nrep} x {L}^2 forward {mcs} MCS")
    t0 = time.perf_counter()
    gpucpm.Batch(sims).run(**kw, iterations=mcs, inner_iterations=mcs)
    for s in sims:
        s.synchronize()
    t_fwd = time.perf_counter() - t0
    p = wls[0].params()
    theta = layout.pack(p["J"], p["lam_area"], p["target_area"], p["lam_perim"], p["target_perim"],
                        dtype=torch.float32).requires_grad_(True)
    gpucpm.energy_stats_many(sims)  # warm
    t0 = time.perf_counter()
    H = energy(theta, sims, layout)
    H.backward()
    t_bwd = time.perf_counter() - t0
    # kernel time: the mean of twenty more launches over the same lattices
    kms = []
    for _ in range(20):
        gpucpm.energy_stats_many(sims)
        kms.append(gpucpm.energy_stats_kernel_ms())
    kernel_ms = sum(kms) / len(kms)
    g = theta.grad.double().numpy()
    lats = []
    for s in sims:
        s.pull_from_gpu()
        lats.append(s.get_state().view(np.uint32).copy())
    H64, t_oracle = _oracle_energy_fn(lats, wls, layout)
    th = theta.detach().double().numpy()
    h = 1e-3
    fd = np.array([(H64(th + h * e) - H64(th - h * e)) / (2 * h) for e in np.eye(th.size)])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1.0)
    nsites = nrep * L * L
    achieved = nsites * 4 / (kernel_ms / 1e3) / 1e9
    out = {"gradient_rel_error": float(rel.max()),
           "energy_rel_error": abs(float(H.detach()) - H64(th)) / abs(H64(th)),
           "replicas": nrep, "L": L, "mcs": mcs, "forward_s": t_fwd,
           "forward_site_updates_per_s": nsites * mcs / t_fwd,
           "backward_s": t_bwd, "backward_kernel_ms": kernel_ms,
           "backward_kernel_ms_runs": kms, "backward_kernel_ms_median": sorted(kms)[len(kms) // 2],
           "backward_sites_per_s": nsites / (kernel_ms / 1e3),
           "roofline": {"bound": "hbm", "kernel": "k_stats_batch<2, unsigned short, true>", "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": _profiled_traffic("k_stats_batch<2, unsigned short, true>", "gradient"),
                        "bytes_per_launch": nsites * 4,
                        "avg_launch_us": kernel_ms * 1e3,
                        "note": "4 algorithmic bytes per site (one int32 read); one launch covers "
                                "all replicas' pair counts and cell moments"},
           "cpu_baseline": {"value": nsites / t_oracle, "unit": "sites/s", "cores": 1, "kind": "port",
                            "sample": f"the oracle's statistics of the same {nrep} lattices, one thread"}}
    if fit:
        from gpucpm.fit import fit_adhesion
        log("contrastive fit of J_12: 16 x 128^2, 500 MCS per evaluation, data at J_12 = 8")
        make = lambda seed: workloads.sorting_small_cells(L=128, seed=seed)

        def replicas(J, lat0, rng0):
            out_ = []
            for r in range(16):
                s = make(lat0 + r).make_cpm()
                s.set_adhesion_real(1, 2, J)
                s.push_to_gpu()
                s.set_seed(rng0 + r)
                out_.append(s)
            return out_

        data = replicas(8.0, 1000, 2000)
        gpucpm.Batch(data).run(**kw, iterations=500, inner_iterations=500)
        t0 = time.perf_counter()
        res = fit_adhesion(data, lambda J: replicas(J, 3000, 4000), (1, 2), 11.0, 500, kw, steps=6)
        out["fit"] = {"parameter": "J_12", "start": 11.0, "data_at": 8.0, "fitted": res.J,
                      "tolerance": 0.5, "within_tolerance": abs(res.J - 8.0) <= 0.5,
                      "evaluations": len(res.history), "seconds": time.perf_counter() - t0,
                      "history": [[float(J), float(gr)] for J, gr in res.history],
                      "workload": "16 replicas of the reference sorting example's Hamiltonian "
                                  "(3x3 cells, area term), 128^2, 500 MCS per evaluation"}
    return out


def measure_interactive(frames, log, no_cpu=False):
    """The notebook pattern (examples/gpucpm_sorting_example.ipynb:316: one
    run(iterations=ticks, inner_iterations=1) per frame, then a readback):
    frames per second on the 256^2 sorting lattice with one MCS and one
    pull_from_gpu per frame, the notebook's shared=0 roles (positions_per_
    thread 4, c 2: 64 colour phases of 1024 attempts per MCS).  The engine
    runs such small phases in one workgroup with the lattice in LDS."""
    from gpucpm import workloads
    wl = workloads.sorting2d(L=256, grid=20, edge=10, seed=2)
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(20240601)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=4, positions_per_checkerboard=2,
              updates_per_checkerboard_switch=1, updates_per_barrier=1,
              iterations=1, inner_iterations=1, shared=0, partial_dispatch=1)
    for _ in range(50):
        sim.run(**kw)
        sim.pull_from_gpu()
    log(f"interactive: {frames} frames of 1 MCS + readback at 256^2")
    t0 = time.perf_counter()
    for _ in range(frames):
        sim.run(**kw)
        sim.pull_from_gpu()
    wall = time.perf_counter() - t0
    sim.timing_begin()
    nt = min(frames, 200)
    for _ in range(nt):
        sim.run(**kw)
    kernel_ms, launches, updates = sim.timing_end()
    launch = sim.last_launch()
    sites = wl.L ** 2
    out = {"workload": "interactive notebook pattern: 2D sorting 256^2, 400 cells, run(iterations=1) + "
                       "pull_from_gpu per frame, shared=0, positions_per_thread 4, c 2",
           "value": frames / wall, "unit": "frames/s", "frames": frames,
           "ms_per_frame": wall / frames * 1e3, "kernel_ms_per_mcs": kernel_ms / nt,
           "launches_per_mcs": launches / nt, "kernel": launch["kernel"], "launch_form": launch["form"],
           "site_updates_per_s_kernel": sites * nt / (kernel_ms / 1e3),
           "cpu_baseline": None}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        p = wl.params()
        prm = O.Params.make(p["temperature"], J=p["J"], S=p["S"], lam_area=p["lam_area"],
                            target_area=p["target_area"], lam_perim=p["lam_perim"],
                            target_perim=p["target_perim"])
        lat = wl.lattice.copy()
        tab = O.tally(lat, 2, wl.ncells)
        cores = len(os.sched_getaffinity(0))
        n = 200
        t0 = time.perf_counter()
        for m in range(n):
            O.sweep(lat, 2, tab, prm, 2, 1, 1, m, 1, cores, shared=False, ppt=4)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "frames/s (MCS/s, no readback)", "cores": cores,
                               "kind": "port", "sample": f"{n} one-MCS sweeps of the same lattice and "
                                                        "schedule, oracle/libcpm_oracle.so with OpenMP"}
    return out


def bench_interactive(a, log):
    r = measure_interactive(a.steps * a.mcs_per_step, log, a.no_cpu)
    line = {"metric": "frames/s (1 MCS + lattice readback per frame)", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": r["ms_per_frame"] * a.mcs_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic", "config": {"workload": r["workload"]}, "roofline": None}
    line.update(r)
    return line


def _profiled_traffic(kernel, workload):
    """HBM bytes per dispatch of kernel on workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/), or None."""
    tfile = os.path.join(ROOT, "profiles", f"traffic_{kernel_token(kernel)}_{workload}.json")
    return json.load(open(tfile))["bytes_per_dispatch"] if os.path.exists(tfile) else None


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
    if a.workload == "interactive":
        print(json.dumps(bench_interactive(a, log)), flush=True)
        return 0
    from gpucpm import workloads

    if a.workload == "batched":
        r = measure_batched(a.replicas, a.steps * a.mcs_per_step, log, a.autotune, a.batch_only)
        line = {"metric": "lattice site-updates/sec (batched replicas)", "n_gpus": 1, "steps": 1,
                "warmup": 1, "ms_per_step": r["ms_total"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
                "config": {"workload": r["workload"]}}
        line.update(r)
        print(json.dumps(line), flush=True)
        return 0
    if a.workload == "gradient":
        r = measure_gradient(32, 512, a.steps * a.mcs_per_step, log, fit=not a.no_grad)
        line = {"metric": "gradient rel. error (dH/dtheta FP32 vs FP64 central differences)",
                "value": r["gradient_rel_error"], "unit": "relative", "n_gpus": 1, "steps": 1,
                "warmup": 0, "ms_per_step": r["forward_s"] * 1e3, "higher_is_better": False,
                "scaling": "weak", "vs_baseline": None,
                "dtype": "int32 sweep, int64 statistics, fp32 gradient", "data": "synthetic",
                "config": {"workload": "32 replicas of 512^2 two-type sorting (C2 parameters), "
                                       "forward then dH/dtheta"}}
        line.update(r)
        print(json.dumps(line), flush=True)
        return 0

    wl = {"sorting2d": workloads.sorting2d, "plain3d": workloads.plain3d,
          "plain2d": workloads.plain2d}[a.workload]()
    c, u = a.checkerboard, a.updates
    r = measure_single(wl, a.workload, c, u, a.steps, a.warmup, a.mcs_per_step, a.cpu_seconds,
                       a.no_cpu, a.autotune, log, a.launch)
    workload_desc = {
        "sorting2d": "2D cell sorting 1024^2, 3969 cells (10x10 squares), two types, 10000-MCS config",
        "plain3d": "3D CPM 256^3, 20000 cubes 5^3, 26-neighbourhood, 1000-MCS config",
        "plain2d": "plain 2D CPM 256^2, 100 cells",
    }[a.workload]
    out = {
        "metric": "lattice site-updates/sec",
        "value": r["value"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "storage": r["storage"],
        "data": "synthetic",
        "config": {"workload": workload_desc, "L": wl.L, "d": wl.d, "cells": wl.ncells,
                   "mcs_per_step": a.mcs_per_step, "positions_per_checkerboard": c,
                   "updates_per_checkerboard_switch": u, "seed": 20240601,
                   "autotune": bool(a.autotune)},
        "mcs_per_s": r["mcs_per_s"],
        "roofline": r["roofline"],
        "cpu_baseline": r["cpu_baseline"],
    }
    if a.workload == "sorting2d" and not a.no_configs:
        # The other configurations, each with its own roofline and baseline.
        configs = {}
        w3 = workloads.plain3d()
        r3 = measure_single(w3, "plain3d", 1, 1, 5, 1, 20, 6.0, a.no_cpu, a.autotune, log)
        r3["workload"] = "3D CPM 256^3, 20000 cubes 5^3, 26-neighbourhood (1000-MCS config; 100 MCS timed)"
        configs["C3_plain3d"] = r3
        # the C2 lattice under the reference's own sorting.py schedule
        # (examples/sorting.py:55-63: shared = 0, positions_per_thread 4, c 2:
        # 64 colour phases per MCS)
        rs = measure_single(workloads.sorting2d(), "sorting2d", 2, 1, 5, 1, 50, 6.0, a.no_cpu, a.autotune,
                            log, shared=0, ppt=4)
        rs["workload"] = ("2D cell sorting 1024^2, 3969 cells, the reference sorting.py schedule: shared=0, "
                          "positions_per_thread 4, positions_per_checkerboard 2 (64 colour phases per MCS)")
        configs["C2_sorting_py_schedule"] = rs
        configs["C4_batched"] = measure_batched(64, 2000, log, a.autotune)
        r5 = measure_gradient(32, 512, 2000, log, fit=not a.no_grad)
        r5["workload"] = ("32 replicas of 512^2 two-type sorting (C2 parameters): forward 2000 MCS "
                          "as one batch, then dH/dtheta for J (3x3), lambda_A, A0, lambda_P, P0 "
                          "against FP64 central differences (step 1e-3)")
        configs["C5_gradient"] = r5
        configs["interactive"] = measure_interactive(2000, log, a.no_cpu)
        out["configs"] = configs
        out["gradient_rel_error"] = r5["gradient_rel_error"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
