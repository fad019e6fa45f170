"""GPU dynamics against the live-read checkerboard (oracle.sweep_live, the
other end of what the reference's concurrent colour phase can produce; see
stats_common) at the reference's sorting.py size, 256^2 over 16 seeds, and
at the benchmark's 1024^2 over 4 seeds: adhesion energy, mean area and
sorting index agree to seed precision (Welch's t within its critical value
at alpha = 1e-4 per comparison, no other tolerance); the area spread is the
snapshot's wider one."""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from helpers import run_kwargs
from stats_common import AREA_STD, compare, observables, sorting_lattice, welch_t

pytestmark = pytest.mark.gpu


def _gpu(L, seed, checkpoints, step):
    import gpucpm
    lat0, n, p, J = sorting_lattice(L=L, seed=seed)
    sim = gpucpm.Cpm(L, 2, n, 10, False)
    sim.set_constraints(cell_type=1, lambda_area=1, target_area=40)
    sim.set_constraints(cell_type=2, lambda_area=1, target_area=40)
    for a, b in ((0, 1), (0, 2), (1, 1), (1, 2), (2, 2)):
        sim.set_constraints(cell_type=a, other_cell_type=b, adhesion=J[a][b])
    sim.get_state()[...] = lat0.view(np.int32)
    sim.push_to_gpu()
    sim.set_seed(77 + seed)
    traj = []
    for _ in range(checkpoints):
        sim.run(**run_kwargs(1, 1, step))
        sim.synchronize()
        sim.pull_from_gpu()
        lat = sim.get_state().view(np.uint32).copy()
        traj.append(observables(lat, n, p, sim.get_cell_table()["area"], J))
    return traj


def _live(L, seed, checkpoints, step):
    lat, n, p, J = sorting_lattice(L=L, seed=seed)
    tab = O.tally(lat, 2, n)
    traj = []
    for k in range(checkpoints):
        O.sweep_live(lat, 2, tab, p, 1, 1, 500 + seed, k * step, step)
        traj.append(observables(lat, n, p, tab.area, J))
    return traj


@pytest.mark.parametrize("L,seeds,checkpoints,step", [(256, 16, 3, 50), (1024, 4, 2, 50)])
def test_gpu_statistics_match_live_read_checkerboard(L, seeds, checkpoints, step):
    print("CASE gpu-stats", L, seeds, flush=True)
    # the oracle's ctypes calls release the GIL: the CPU runs overlap the GPU's
    with ThreadPoolExecutor(min(seeds, 16)) as ex:
        live = ex.map(lambda s: _live(L, s, checkpoints, step), range(seeds))
        gpu = np.array([_gpu(L, s, checkpoints, step) for s in range(seeds)])
        live = np.array(list(live))
    ok, rows = compare(gpu, live, skip=(AREA_STD,))
    for r in rows:
        print("t=%d %-14s gpu %.5g live %.5g t=%.2f (critical %.2f)" % ((r[0] + 1) * step, *r[1:]))
    assert ok, [r for r in rows if abs(r[4]) > r[5]]
    g, lv = gpu[:, :, AREA_STD].mean(1), live[:, :, AREA_STD].mean(1)
    t, crit = welch_t(g, lv)
    print("area std gpu %.4f live %.4f (ratio %.3f) t=%.2f" % (g.mean(), lv.mean(), g.mean() / lv.mean(), t))
    assert t > -crit  # never narrower than under live reads
    assert gpu[:, -1, 1].mean() > 30
